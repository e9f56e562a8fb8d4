// Host-side internals of the engine: the handle, device buffers, error
// plumbing and the kernel launch wrappers shared by the C-ABI translation
// units.
#pragma once
#include <array>
#include <map>
#include <utility>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "fk_api.h"

namespace fk {

// Thrown inside the library, converted to an fk_status at the C boundary.
struct Error : std::runtime_error {
    fk_status code;
    Error(fk_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define FK_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::fk::Error(e_ == hipErrorOutOfMemory ? FK_E_NOMEM : FK_E_DEVICE,       \
                              std::string(#call) + ": " + hipGetErrorString(e_));         \
    } while (0)

#define FK_CHECK_ARG(cond, msg)                                                           \
    do {                                                                                  \
        if (!(cond)) throw ::fk::Error(FK_E_INVALID, msg);                                \
    } while (0)

#define FK_LAUNCH_CHECK() FK_HIP(hipGetLastError())

// Device memory cache of one handle.  Blocks its buffers release are kept
// and handed out again for requests they hold with at most a quarter to
// spare; they go back to HIP when the handle closes, when an allocation
// finds the device full, or when the cache would hold more than a quarter of
// the device.  A build allocates tens of buffers (pools, lists, candidate
// buckets of a few GB each) and hipMalloc / hipFree of such sizes cost
// milliseconds each, and hipFree waits for the device; with the cache a
// repeated build allocates nothing.  The handle's work runs in order on its
// one stream, so a block reused by the handle cannot be touched by an
// earlier owner's launch after the new owner's first one.
struct MemCache {
    struct Blk {
        void *p;
        size_t bytes;
    };
    std::vector<Blk> free_;
    size_t held = 0, limit = 0;
    void *get(size_t bytes, size_t &got);
    void put(void *p, size_t bytes);
    void trim();
    ~MemCache() { trim(); }
};
// The cache of the handle whose C-ABI call runs on this thread (nullptr:
// plain hipMalloc / hipFree).
MemCache *&current_cache();

// Owning device allocation; freed (or returned to its handle's cache) on
// scope exit.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    MemCache *mc = nullptr;  // the cache the block came from
    size_t blk = 0;          // its size in bytes
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), mc(o.mc), blk(o.blk) { o.p = nullptr; o.n = 0; o.mc = nullptr; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; mc = o.mc; blk = o.blk;
            o.p = nullptr; o.n = 0; o.mc = nullptr;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        mc = current_cache();
        if (mc) {
            p = static_cast<T *>(mc->get(count * sizeof(T), blk));
        } else {
            FK_HIP(hipMalloc((void **)&p, count * sizeof(T)));
            blk = count * sizeof(T);
        }
        n = count;
    }
    // Grow-only reallocation (contents discarded); a buffer that grows
    // again gets a quarter more than asked.
    void ensure(size_t count) {
        if (p == nullptr) alloc(count);
        else if (count > n) alloc(count + count / 4);
    }
    void release() {
        if (p) {
            if (mc) mc->put(p, blk);
            else (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        mc = nullptr;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Striped device counters.  Atomics to one address are performed one at a
// time, and a launch whose million waves each add a count to the same word
// stalls the memory pipeline of every CU behind that word.  A tally counter
// is FK_TALLY_STRIPES words 128 bytes apart: a workgroup adds to the stripe
// of its index, and the host sums the stripes after the launch.
#define FK_TALLY_STRIPES 64
#define FK_TALLY_STRIDE 16  // words per stripe (128 bytes)
#define FK_TALLY_WORDS (FK_TALLY_STRIPES * FK_TALLY_STRIDE)

__device__ __forceinline__ void tally_add(unsigned long long *t, unsigned long long v) {
    atomicAdd(t + (blockIdx.x & (FK_TALLY_STRIPES - 1)) * FK_TALLY_STRIDE, v);
}

// Host side: `slots` counters of FK_TALLY_WORDS words each.
struct Tally {
    DevBuf<unsigned long long> buf;
    std::vector<unsigned long long> host;
    void ensure(size_t slots) { buf.ensure(slots * FK_TALLY_WORDS); }
    unsigned long long *at(size_t k) const { return buf.p + k * FK_TALLY_WORDS; }
    void clear(hipStream_t s, size_t k, size_t n = 1) {
        FK_HIP(hipMemsetAsync(at(k), 0, n * FK_TALLY_WORDS * sizeof(unsigned long long), s));
    }
    // Sums of counters [k, k + n) into out[0, n); synchronizes the stream.
    void read(hipStream_t s, size_t k, size_t n, unsigned long long *out) {
        host.resize(n * FK_TALLY_WORDS);
        FK_HIP(hipMemcpyAsync(host.data(), at(k), n * FK_TALLY_WORDS * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, s));
        FK_HIP(hipStreamSynchronize(s));
        for (size_t c = 0; c < n; ++c) {
            unsigned long long v = 0;
            for (size_t j = 0; j < FK_TALLY_STRIPES; ++j) v += host[c * FK_TALLY_WORDS + j * FK_TALLY_STRIDE];
            out[c] = v;
        }
    }
    unsigned long long read1(hipStream_t s, size_t k) {
        unsigned long long v = 0;
        read(s, k, 1, &v);
        return v;
    }
};

// Pinned (page-locked) host buffer, grow-only.
template <typename T>
struct PinnedBuf {
    T *p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    void ensure(size_t count) {
        if (count <= n && p) return;
        release();
        FK_HIP(hipHostMalloc((void **)&p, (count ? count : 1) * sizeof(T), hipHostMallocDefault));
        n = count;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    ~PinnedBuf() { release(); }
};

struct Dataset {
    DevBuf<float> rows;   // n_total x stride floats, zero padded
    DevBuf<float> chains; // n_total x dpad floats, chain-major copy (octet distances)
    DevBuf<float> center; // dpad floats: Calcenter of all n_total rows (float row-order sums), zero padded
    uint32_t n_total = 0;
    uint32_t d = 0;
    uint32_t dpad = 0;    // d rounded up to a multiple of 8
    uint32_t stride = 0;  // floats between rows (multiple of 8)
    // Gram-form join (fk_gram.hip): per-row (|x~|^2, |x|), operand format
    // (FK_GRAM_*), FP16 prescale and the FK_FLAG_GRAM_* request it was made for
    DevBuf<float2> gram_nrm;
    DevBuf<uint8_t> chains8;  // exact form with values in [0, 255]: chain-major byte copy
    DevBuf<int32_t> norms8;   // with chains8: sum over dpad of (byte - 128)^2 per row
    DevBuf<uint4> rows8q;     // widths 96 / 256, FP16 form: per row 128 / 384 bytes: int8 q[dpad], FP32 scale s,
                              // error norm |x - s q| rounded up, |s q|^2 (BridgeView bound pass)
    DevBuf<uint16_t> rows16;  // n_total x dpad 16-bit Gram operands (FP16 of the prescaled value, else BF16)
    int gram_fmt = 0;
    float gram_scale = 1.f;
    float gram_ymax = 0.f;  // largest row norm |x| over the rows, rounded up
    uint32_t gram_req = 0;
    // centroids of row ranges (batched builds' sets) already made, d floats
    // each, keyed by (first row, row count); cleared by every upload
    mutable std::map<std::pair<uint32_t, uint32_t>, std::vector<float>> range_center;
};

}  // namespace fk

struct fk_handle {
    fk::MemCache cache;  // first member: destroyed after every buffer below
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    fk::Dataset ds;
    fk::Tally counters;  // kernel counters (tally slots)
    fk::PinnedBuf<uint32_t> host_rows;         // connectivity pass: pool rows of attached ids
    fk::PinnedBuf<uint32_t> host_lab, host_deg;  // connectivity pass: labels and out-degrees
    fk::PinnedBuf<uint8_t> dl_stage[2];        // device-to-host staging (download_bytes)
    uint8_t *noalloc = nullptr;                // device address of the fk_prune_noalloc scratch array
    // prune kernel variant chosen per shape by measurement (fk_build.hip:
    // tuned_prune): key = rule, dpad, row kind, nsg_R, maxc, reversed
    std::map<std::array<uint32_t, 6>, int> prune_variant;
};

namespace fk {

// Launch wrappers implemented in the kernel translation units.  All launch on
// the handle's stream and take device pointers.
void launch_pair_distances(hipStream_t s, const Dataset &ds, const uint32_t *a, const uint32_t *b,
                           size_t n, float *out);

struct PruneLaunch {
    const uint64_t *cand = nullptr;   // per-task sorted candidate keys
    const uint32_t *cand_len = nullptr;
    uint64_t cand_stride = 0;         // fixed stride, used when cand_off is null
    const uint64_t *cand_off = nullptr;  // optional CSR offsets per slot
    const uint32_t *qid = nullptr;    // owner node per task (nullptr: task index)
    const uint32_t *task_list = nullptr;  // optional indirection: task -> pool slot
    uint32_t task_base = 0;           // the launch covers tasks [task_base, task_base + n_tasks)
    bool xcd_order = false;           // map blocks to tasks XCD-contiguously
    uint32_t n_tasks = 0;
    uint32_t oracle_n = 0;            // ids >= oracle_n are skipped
    uint32_t maxc = 0;                // examined positions < min(len, maxc)
    uint32_t R = 0;
    float tau3 = 0.f, cos_thr = 0.f;
    int rule = 0;
    uint32_t *out_ids = nullptr;
    float *out_d = nullptr;
    uint64_t out_stride = 0;
    uint32_t *out_len = nullptr;
    unsigned long long *cc = nullptr;  // accumulated distance count (tally counter)
    unsigned long long *examined = nullptr;  // accumulated examined candidates (tally counter)
    unsigned long long *uncomputed = nullptr;  // counted distances not computed (tau rule, pd <= 3 tau)
    int variant = 0;                  // kernel variant (FK_PRUNE_VARIANTS; same results)
    bool gram = false;                // variant 3 may use the dataset's Gram preparation
    const float2 *nrm = nullptr;      // exact-form rows: norms for dot-product distances
    const uint8_t *chains8 = nullptr; // exact-form byte rows (chain-major), with nrm
};
void launch_prune(hipStream_t s, const Dataset &ds, const PruneLaunch &a);
#define FK_PRUNE_VARIANTS 4
// Whether variant v runs code of its own for the launch's shape (else it
// runs as variant 0).
bool prune_variant_applies(const Dataset &ds, const PruneLaunch &a, int v);
// Prune variant 3 (fk_prune_gram.hip): applies to a.gram launches on a
// Gram-prepared dataset with R <= 64; launch_prune_gram returns false, and
// launches nothing, where it does not.
bool prune_gram_applies(const Dataset &ds, const PruneLaunch &a);
bool launch_prune_gram(hipStream_t s, const Dataset &ds, const PruneLaunch &a);
// Whether a build's local join takes the Gram form: FK_FLAG_GRAM on exact-
// form rows, or with a bounded form forced by FK_FLAG_GRAM_BF16 / _F16.  On
// the bounded forms a pair that enters a pool needs its canonical FP32
// distance (its key) anyway and a third to a half of the joined pairs do, so
// the plain join is the faster one unless a bounded form is asked for.
inline bool join_in_gram_form(uint32_t flags, const Dataset &ds) {
    return (flags & FK_FLAG_GRAM) && (ds.gram_fmt == FK_GRAM_EXACT || (flags & (FK_FLAG_GRAM_BF16 | FK_FLAG_GRAM_F16)));
}
// Adds an order-independent digest of the launch's output lists to *out.
void launch_prune_digest(hipStream_t s, const PruneLaunch &a, unsigned long long *out);
void build_chains(hipStream_t s, Dataset &ds);  // fills ds.chains from ds.rows

void launch_join_distances(hipStream_t s, const Dataset &ds, const uint32_t *ids,
                           const uint64_t *list_off, const uint32_t *n_new, const uint64_t *pair_off,
                           uint32_t n_nodes, uint32_t tmax, float *out);

// Gram-form distances (fk_gram.hip): prepare the dataset for the flags'
// operand format (cached), and pair distances with their bounds.
void gram_prepare(hipStream_t s, Dataset &ds, uint32_t flags);
void launch_gram_pairs(hipStream_t s, const Dataset &ds, const uint32_t *a, const uint32_t *b, size_t n,
                       float *out, float *lo, float *hi);

// Exclusive prefix sums over device arrays (out may alias in).  `total`
// (device pointer, may be nullptr) receives the sum; `tmp` is grown as needed
// and must stay alive until the stream has run the scan.
void exclusive_scan_u32_to_u64(hipStream_t s, const uint32_t *in, uint64_t *out, size_t n,
                               uint64_t *total, DevBuf<uint64_t> &tmp);
void exclusive_scan_u64(hipStream_t s, const uint64_t *in, uint64_t *out, size_t n, uint64_t *total,
                        DevBuf<uint64_t> &tmp);

// Brute-force exact kNN of device queries against rows [base, base + n_used);
// d_self holds ids relative to base, the output ids are absolute.
void bruteforce_knn(hipStream_t s, const Dataset &ds, const float *d_queries, uint32_t q_stride,
                    uint32_t n_q, uint32_t n_used, uint32_t K, const uint32_t *d_self,
                    uint32_t *d_out_ids, float *d_out_dists, uint32_t base = 0);

}  // namespace fk
