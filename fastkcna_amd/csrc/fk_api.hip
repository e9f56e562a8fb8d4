// C-ABI entry points: handle lifetime, dataset residency and the unit-level
// entry points (distances, prune, brute force).  The full build and
// NN-descent live in fk_build.hip, the query search in fk_search.hip.
#include <math.h>

#include <sys/mman.h>

#include <algorithm>
#include <new>
#include <mutex>
#include <thread>

#include "fk_api_impl.h"
#include "fk_device.h"
#include "fk_internal.h"

namespace fk {

float alpha_cos_threshold(float angle) {
    // src/kgraph.cpp:3160-3161: float angle; std::cos(angle / 180 * kPi) with
    // kPi a double constant; the float quotient is widened, cos is taken in
    // double and the result narrowed to float.
    const double kPi = 3.14159265358979323846264;
    return (float)cos((double)(angle / 180.0f) * kPi);
}

int rule_of(fk_pg_type t) {
    switch (t) {
        case FK_PG_TAUMNG: return FK_RULE_TAU;
        case FK_PG_ALPHAPG: return FK_RULE_ALPHA;
        default: return FK_RULE_NSG;
    }
}

void select_device(fk_handle *h) { FK_HIP(hipSetDevice(h->device)); }

// Host rows (any stride) -> device rows of dpad floats, zero padded, through
// two pinned staging buffers so packing and PCIe transfer overlap.  With
// colsum, the rows are also added into colsum[0, d) in float, one row after
// the other (the summation order of MatrixOracle::Calcenter,
// src/kgraph-data.h:413-428).
void upload_rows(fk_handle *h, float *dst, uint32_t dpad, const float *src, size_t src_stride_bytes, uint32_t n,
                 uint32_t d, float *colsum) {
    const size_t row_bytes = (size_t)dpad * sizeof(float);
    const size_t chunk_rows = std::max<size_t>(1, ((size_t)32 << 20) / row_bytes);
    float *stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    struct Cleanup {
        float **s;
        hipEvent_t *e;
        ~Cleanup() {
            for (int i = 0; i < 2; ++i) {
                if (e[i]) { (void)hipEventSynchronize(e[i]); (void)hipEventDestroy(e[i]); }
                if (s[i]) (void)hipHostFree(s[i]);
            }
        }
    } cleanup{stage, done};
    for (int i = 0; i < 2; ++i) {
        FK_HIP(hipHostMalloc((void **)&stage[i], chunk_rows * row_bytes, hipHostMallocDefault));
        FK_HIP(hipEventCreate(&done[i]));
    }
    int k = 0;
    for (size_t r0 = 0; r0 < n; r0 += chunk_rows, k ^= 1) {
        const size_t m = std::min<size_t>(chunk_rows, n - r0);
        FK_HIP(hipEventSynchronize(done[k]));  // the buffer's previous copy has finished
        for (size_t r = 0; r < m; ++r) {
            float *o = stage[k] + r * dpad;
            memcpy(o, (const char *)src + (r0 + r) * src_stride_bytes, (size_t)d * sizeof(float));
            if (dpad != d) memset(o + d, 0, (size_t)(dpad - d) * sizeof(float));
            if (colsum)
                for (uint32_t c = 0; c < d; ++c) colsum[c] += o[c];
        }
        FK_HIP(hipMemcpyAsync(dst + r0 * dpad, stage[k], m * row_bytes, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipEventRecord(done[k], h->stream));
    }
    FK_HIP(hipStreamSynchronize(h->stream));
}

MemCache *&current_cache() {
    static thread_local MemCache *c = nullptr;
    return c;
}

static size_t cache_round(size_t bytes) {
    const size_t g = bytes <= (1u << 20) ? 512 : (2u << 20);
    return (bytes + g - 1) / g * g;
}

void *MemCache::get(size_t bytes, size_t &got) {
    const size_t want = cache_round(bytes);
    size_t best = free_.size();
    for (size_t i = 0; i < free_.size(); ++i) {
        const size_t b = free_[i].bytes;
        if (b >= want && b <= want + want / 4 && (best == free_.size() || b < free_[best].bytes)) best = i;
    }
    if (best < free_.size()) {
        void *p = free_[best].p;
        got = free_[best].bytes;
        held -= got;
        free_[best] = free_.back();
        free_.pop_back();
        return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipErrorOutOfMemory && !free_.empty()) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(e == hipErrorOutOfMemory ? FK_E_NOMEM : FK_E_DEVICE,
                    "hipMalloc(" + std::to_string(want) + "): " + hipGetErrorString(e));
    }
    got = want;
    return p;
}

void MemCache::put(void *p, size_t bytes) {
    if (limit == 0) {
        size_t fr = 0, tot = 0;
        limit = hipMemGetInfo(&fr, &tot) == hipSuccess ? tot / 4 : ((size_t)16 << 30);
    }
    if (held + bytes > limit) {
        (void)hipFree(p);
        return;
    }
    free_.push_back({p, bytes});
    held += bytes;
}

void MemCache::trim() {
    for (const Blk &b : free_) (void)hipFree(b.p);
    free_.clear();
    held = 0;
}

// Host memory for a result the caller frees with free(): 2 MB aligned and
// advised for huge pages, so that filling it takes few page faults.
void *host_result_alloc(size_t bytes) {
    const size_t huge = (size_t)2 << 20;
    const size_t n = (std::max<size_t>(bytes, 1) + huge - 1) / huge * huge;
    void *p = aligned_alloc(huge, n);
    if (p && n >= huge) (void)madvise(p, n, MADV_HUGEPAGE);
    return p;
}

// memcpy split over FK_COPY_THREADS threads (the host copy out of pinned
// staging is what bounds a download).
#define FK_COPY_THREADS 4
static void copy_parallel(char *dst, const char *src, size_t n) {
    if (n < ((size_t)1 << 20)) {
        memcpy(dst, src, n);
        return;
    }
    const size_t part = (n / FK_COPY_THREADS + 63) & ~(size_t)63;
    std::thread t[FK_COPY_THREADS - 1];
    for (int i = 1; i < FK_COPY_THREADS; ++i) {
        const size_t a = std::min(n, (size_t)i * part), e = std::min(n, a + part);
        t[i - 1] = std::thread([=] { memcpy(dst + a, src + a, e - a); });
    }
    memcpy(dst, src, std::min(n, part));
    for (auto &x : t) x.join();
}

// Copies `bytes` of device memory to host memory through the handle's two
// pinned 16 MB staging buffers: the DMA of one chunk overlaps the host copy
// of the one before.  Synchronizes the handle's stream.
void download_bytes(fk_handle *h, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    const size_t chunk = (size_t)16 << 20;
    for (int i = 0; i < 2; ++i) h->dl_stage[i].ensure(chunk);
    hipEvent_t ev[2] = {nullptr, nullptr};
    struct Events {
        hipEvent_t *e;
        ~Events() {
            for (int i = 0; i < 2; ++i)
                if (e[i]) { (void)hipEventSynchronize(e[i]); (void)hipEventDestroy(e[i]); }
        }
    } events{ev};
    for (int i = 0; i < 2; ++i) FK_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    const size_t nch = (bytes + chunk - 1) / chunk;
    for (size_t i = 0; i <= nch; ++i) {
        if (i < nch) {  // chunk i into the stage chunk i - 2 left (copied out last step)
            const size_t m = std::min(chunk, bytes - i * chunk);
            FK_HIP(hipMemcpyAsync(h->dl_stage[i & 1].p, (const char *)src + i * chunk, m, hipMemcpyDeviceToHost,
                                  h->stream));
            FK_HIP(hipEventRecord(ev[i & 1], h->stream));
        }
        if (i >= 1) {
            const size_t j = i - 1;
            const size_t m = std::min(chunk, bytes - j * chunk);
            FK_HIP(hipEventSynchronize(ev[j & 1]));
            copy_parallel((char *)dst + j * chunk, (const char *)h->dl_stage[j & 1].p, m);
        }
    }
}

void require_dataset(fk_handle *h) {
    if (h->ds.n_total == 0) throw Error(FK_E_INVALID, "no dataset uploaded (fk_dataset_upload)");
}

// ---------------------------------------------------------------------------
// fk_prune_noalloc: the prune with every device buffer in `scratch`, a
// file-scope device array (each GPU holds its own copy from the first use of
// the library on it until the process ends).  Host code never passes the
// array's name to a kernel: it uses the device address hipGetSymbolAddress
// returns, looked up once per device by fk_open (noalloc_scratch) and kept in
// the handle.  Layout of one call, each part 256-byte aligned: the distance
// tally (FK_TALLY_WORDS u64), pools, pool lengths, owners, output ids,
// output distances, output lengths.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint8_t scratch[FK_PRUNE_NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on `device`, which must be current (one lookup
// per device, kept for the process).
uint8_t *noalloc_scratch(int device) {
    static std::mutex mu;
    static std::vector<uint8_t *> addr;
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)device >= addr.size()) addr.resize((size_t)device + 1, nullptr);
    if (!addr[device]) {
        void *p = nullptr;
        FK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(FK_E_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
        addr[device] = static_cast<uint8_t *>(p);
    }
    return addr[device];
}

// Zero fill of n u64 words (in place of a memset node, so that the call has
// only kernels and copies).
__global__ void k_fill_zero_u64(unsigned long long *__restrict__ p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of scratch a call of n_tasks pools needs.
size_t noalloc_bytes(uint32_t n_tasks, uint32_t pool_stride, uint32_t out_stride) {
    const size_t n = n_tasks;
    return align256((size_t)FK_TALLY_WORDS * 8) + align256(n * pool_stride * 8) + 2 * align256(n * 4) +
           2 * align256(n * out_stride * 4) + align256(n * 4);
}

}  // namespace fk

using namespace fk;

extern "C" {

const char *fk_version(void) { return "fastkcna-amd-0.1 gfx950"; }

fk_status fk_open(int device, fk_handle **out) {
    if (!out) return FK_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device < 0 || device >= n) return FK_E_NODEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return FK_E_NODEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return FK_E_NODEVICE;
    fk_handle *h = new (std::nothrow) fk_handle();
    if (!h) return FK_E_NOMEM;
    h->device = device;
    fk_status st = guard(h, [&] {
        select_device(h);
        FK_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->counters.ensure(2);
        h->noalloc = noalloc_scratch(device);  // this device's copy of the fk_prune_noalloc scratch
    });
    if (st != FK_OK) {
        delete h;
        return st;
    }
    *out = h;
    return FK_OK;
}

void fk_close(fk_handle *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->ds.rows.release();
    h->ds.chains.release();
    h->ds.center.release();
    h->counters.buf.release();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *fk_last_error(const fk_handle *h) { return h ? h->last_error.c_str() : "null handle"; }

fk_status fk_dataset_upload(fk_handle *h, const float *rows, size_t row_stride_bytes, uint32_t n_total,
                            uint32_t d) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        FK_CHECK_ARG(rows != nullptr && n_total > 0 && d > 0, "rows must be non-empty");
        FK_CHECK_ARG(row_stride_bytes >= (size_t)d * sizeof(float), "row stride smaller than a row");
        select_device(h);
        Dataset &ds = h->ds;
        ds.rows.release();
        ds.chains.release();
        ds.center.release();
        ds.gram_nrm.release();
        ds.chains8.release();
        ds.norms8.release();
        ds.rows16.release();
        ds.rows8q.release();
        ds.gram_fmt = FK_GRAM_NONE;
        ds.gram_req = 0;
        ds.n_total = 0;
        ds.range_center.clear();
        const uint32_t dpad = (d + 7u) & ~7u;
        ds.rows.alloc((size_t)n_total * dpad);
        // the centroid of all rows (Calcenter: float sums in row order, then
        // a float division by the row count), zero padded to dpad
        std::vector<float> mean(dpad, 0.f);
        upload_rows(h, ds.rows.p, dpad, rows, row_stride_bytes, n_total, d, mean.data());
        const float count = (float)n_total;
        for (uint32_t c = 0; c < d; ++c) mean[c] /= count;
        ds.center.alloc(dpad);
        FK_HIP(hipMemcpyAsync(ds.center.p, mean.data(), (size_t)dpad * 4, hipMemcpyHostToDevice, h->stream));
        ds.n_total = n_total;
        ds.d = d;
        ds.dpad = dpad;
        ds.stride = dpad;
        build_chains(h->stream, ds);
        FK_HIP(hipStreamSynchronize(h->stream));
    });
}

fk_status fk_pair_distances(fk_handle *h, const uint32_t *a, const uint32_t *b, size_t n, float *out) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        FK_CHECK_ARG(n == 0 || (a && b && out), "null pointer");
        for (size_t i = 0; i < n; ++i)
            FK_CHECK_ARG(a[i] < h->ds.n_total && b[i] < h->ds.n_total, "row id out of range");
        if (n == 0) return;
        select_device(h);
        DevBuf<uint32_t> da(n), db(n);
        DevBuf<float> dout(n);
        FK_HIP(hipMemcpyAsync(da.p, a, n * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(db.p, b, n * 4, hipMemcpyHostToDevice, h->stream));
        launch_pair_distances(h->stream, h->ds, da.p, db.p, n, dout.p);
        FK_HIP(hipMemcpyAsync(out, dout.p, n * 4, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipStreamSynchronize(h->stream));
    });
}

fk_status fk_gram_pair_distances(fk_handle *h, uint32_t flags, const uint32_t *a, const uint32_t *b, size_t n,
                                 float *out, float *lo, float *hi, uint32_t *mode) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        FK_CHECK_ARG(n == 0 || (a && b && out && lo && hi), "null pointer");
        FK_CHECK_ARG(!((flags & FK_FLAG_GRAM_BF16) && (flags & FK_FLAG_GRAM_F16)), "conflicting Gram formats");
        for (size_t i = 0; i < n; ++i)
            FK_CHECK_ARG(a[i] < h->ds.n_total && b[i] < h->ds.n_total, "row id out of range");
        select_device(h);
        gram_prepare(h->stream, h->ds, flags);
        if (mode) *mode = (uint32_t)h->ds.gram_fmt;
        if (n == 0) return;
        DevBuf<uint32_t> da(n), db(n);
        DevBuf<float> dout(3 * n);
        FK_HIP(hipMemcpyAsync(da.p, a, n * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(db.p, b, n * 4, hipMemcpyHostToDevice, h->stream));
        launch_gram_pairs(h->stream, h->ds, da.p, db.p, n, dout.p, dout.p + n, dout.p + 2 * n);
        FK_HIP(hipMemcpyAsync(out, dout.p, n * 4, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipMemcpyAsync(lo, dout.p + n, n * 4, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipMemcpyAsync(hi, dout.p + 2 * n, n * 4, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipStreamSynchronize(h->stream));
    });
}

fk_status fk_join_distances(fk_handle *h, const uint32_t *ids, const uint64_t *list_off, const uint32_t *n_new,
                            uint32_t n_nodes, float *out, uint64_t *n_pairs) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        FK_CHECK_ARG(list_off && n_new, "null pointer");
        const uint64_t n_ids = list_off[n_nodes];
        FK_CHECK_ARG(n_ids == 0 || ids, "null ids");
        std::vector<uint64_t> pair_off(n_nodes + 1, 0);
        uint64_t tmax = 0;
        for (uint32_t t = 0; t < n_nodes; ++t) {
            FK_CHECK_ARG(list_off[t + 1] >= list_off[t], "list offsets must be non-decreasing");
            const uint64_t T = list_off[t + 1] - list_off[t];
            tmax = std::max(tmax, T);
            const uint64_t nn = n_new[t];
            FK_CHECK_ARG(nn <= T, "n_new larger than the list");
            pair_off[t + 1] = pair_off[t] + nn * (nn - (nn ? 1 : 0)) / 2 + nn * (T - nn);
        }
        for (uint64_t i = 0; i < n_ids; ++i) FK_CHECK_ARG(ids[i] < h->ds.n_total, "id out of range");
        if (n_pairs) *n_pairs = pair_off[n_nodes];
        if (pair_off[n_nodes] == 0 || n_nodes == 0) return;
        FK_CHECK_ARG(out, "null output");
        select_device(h);
        DevBuf<uint32_t> dids(n_ids), dnn(n_nodes);
        DevBuf<uint64_t> dlo(n_nodes + 1), dpo(n_nodes + 1);
        DevBuf<float> dout(pair_off[n_nodes]);
        FK_HIP(hipMemcpyAsync(dids.p, ids, n_ids * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(dnn.p, n_new, (size_t)n_nodes * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(dlo.p, list_off, (size_t)(n_nodes + 1) * 8, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(dpo.p, pair_off.data(), (size_t)(n_nodes + 1) * 8, hipMemcpyHostToDevice,
                              h->stream));
        FK_CHECK_ARG(tmax <= 16384, "sample list longer than 16384");
        launch_join_distances(h->stream, h->ds, dids.p, dlo.p, dnn.p, dpo.p, n_nodes, (uint32_t)tmax, dout.p);
        FK_HIP(hipMemcpyAsync(out, dout.p, pair_off[n_nodes] * 4, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipStreamSynchronize(h->stream));
    });
}

// Argument checks shared by fk_prune and fk_prune_noalloc.
static void prune_checks(const fk_handle *h, const uint64_t *pools, const uint32_t *pool_len, uint32_t pool_stride,
                         const uint32_t *q, uint32_t n_tasks, const fk_params *p, int reversed,
                         const uint32_t *out_ids, uint32_t out_stride, const uint32_t *out_len) {
    FK_CHECK_ARG(p && pools && pool_len && out_ids && out_len, "null pointer");
    FK_CHECK_ARG(out_stride >= p->nsg_R, "out_stride < nsg_R");
    FK_CHECK_ARG(p->nsg_R <= 256, "nsg_R above 256 is not supported");
    for (uint32_t t = 0; t < n_tasks; ++t) {
        FK_CHECK_ARG(pool_len[t] <= pool_stride, "pool_len > pool_stride");
        const uint64_t *c = pools + (uint64_t)t * pool_stride;
        // Rows the pruner reads must be resident: pool[start] is always
        // kept; the sync form skips later ids >= oracle_size by itself,
        // the reversed form reads every entry.
        const uint32_t qt = q ? q[t] : t;
        const uint32_t start = (pool_len[t] > 0 && (uint32_t)c[0] == qt) ? 1u : 0u;
        if (start < pool_len[t]) FK_CHECK_ARG((uint32_t)c[start] < h->ds.n_total, "pool id out of range");
        if (reversed)
            for (uint32_t k = 0; k < pool_len[t]; ++k)
                FK_CHECK_ARG((uint32_t)c[k] < h->ds.n_total, "pool id out of range");
    }
}

// The launch arguments of one prune over device-resident pools (everything
// but the buffers).
static PruneLaunch prune_args(const fk_handle *h, fk_pg_type type, uint32_t pool_stride, uint32_t n_tasks,
                              uint32_t oracle_size, const fk_params *p, int reversed, uint32_t out_stride) {
    PruneLaunch a;
    if ((p->flags & FK_FLAG_GRAM) && h->ds.gram_fmt == FK_GRAM_EXACT) {
        // exact-form rows: occlusion distances from dot products
        a.nrm = h->ds.gram_nrm.p;
        a.chains8 = h->ds.chains8.p;
    }
    const uint32_t lim = oracle_size < h->ds.n_total ? oracle_size : h->ds.n_total;
    a.cand_stride = pool_stride;
    a.n_tasks = n_tasks;
    a.oracle_n = reversed ? 0xFFFFFFFFu : lim;
    a.maxc = reversed ? 0xFFFFFFFFu : p->search_K;
    a.R = p->nsg_R;
    a.tau3 = 3.0f * p->tau;
    a.cos_thr = alpha_cos_threshold(p->angle);
    a.rule = rule_of(type);
    a.out_stride = out_stride;
    a.variant = (p->flags & FK_FLAG_PRUNE_V1) ? 1 : (p->flags & FK_FLAG_PRUNE_V2) ? 2 : (p->flags & FK_FLAG_PRUNE_V3) ? 3 : 0;
    a.gram = (p->flags & FK_FLAG_GRAM) != 0;
    return a;
}

fk_status fk_prune(fk_handle *h, fk_pg_type type, const uint64_t *pools, const uint32_t *pool_len,
                   uint32_t pool_stride, const uint32_t *q, uint32_t n_tasks, uint32_t oracle_size,
                   const fk_params *p, int reversed, uint32_t *out_ids, float *out_dists, uint32_t out_stride,
                   uint32_t *out_len, uint64_t *cc) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        prune_checks(h, pools, pool_len, pool_stride, q, n_tasks, p, reversed, out_ids, out_stride, out_len);
        if (n_tasks == 0) {
            if (cc) *cc = 0;
            return;
        }
        select_device(h);
        const size_t pool_n = (size_t)n_tasks * pool_stride;
        DevBuf<uint64_t> dpool(pool_n);
        DevBuf<uint32_t> dlen(n_tasks), dq(n_tasks), doid((size_t)n_tasks * out_stride), dolen(n_tasks);
        DevBuf<float> dod((size_t)n_tasks * out_stride);
        FK_HIP(hipMemcpyAsync(dpool.p, pools, pool_n * 8, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(dlen.p, pool_len, (size_t)n_tasks * 4, hipMemcpyHostToDevice, h->stream));
        if (q) FK_HIP(hipMemcpyAsync(dq.p, q, (size_t)n_tasks * 4, hipMemcpyHostToDevice, h->stream));
        h->counters.clear(h->stream, 0);
        if (p->flags & FK_FLAG_GRAM) gram_prepare(h->stream, h->ds, p->flags);
        PruneLaunch a = prune_args(h, type, pool_stride, n_tasks, oracle_size, p, reversed, out_stride);
        a.cand = dpool.p;
        a.cand_len = dlen.p;
        a.qid = q ? dq.p : nullptr;
        a.out_ids = doid.p;
        a.out_d = dod.p;
        a.out_len = dolen.p;
        a.cc = h->counters.at(0);
        launch_prune(h->stream, h->ds, a);
        FK_HIP(hipMemcpyAsync(out_ids, doid.p, (size_t)n_tasks * out_stride * 4, hipMemcpyDeviceToHost,
                              h->stream));
        if (out_dists)
            FK_HIP(hipMemcpyAsync(out_dists, dod.p, (size_t)n_tasks * out_stride * 4, hipMemcpyDeviceToHost,
                                  h->stream));
        FK_HIP(hipMemcpyAsync(out_len, dolen.p, (size_t)n_tasks * 4, hipMemcpyDeviceToHost, h->stream));
        const unsigned long long c = h->counters.read1(h->stream, 0);
        if (cc) *cc = c;
    });
}

fk_status fk_prune_noalloc(fk_handle *h, fk_pg_type type, const uint64_t *pools, const uint32_t *pool_len,
                           uint32_t pool_stride, const uint32_t *q, uint32_t n_tasks, uint32_t oracle_size,
                           const fk_params *p, int reversed, uint32_t *out_ids, float *out_dists,
                           uint32_t out_stride, uint32_t *out_len, uint64_t *cc) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        prune_checks(h, pools, pool_len, pool_stride, q, n_tasks, p, reversed, out_ids, out_stride, out_len);
        const size_t need = noalloc_bytes(n_tasks, pool_stride, out_stride);
        if (need > FK_PRUNE_NOALLOC_SCRATCH_BYTES)
            throw Error(FK_E_INVALID, "fk_prune_noalloc: " + std::to_string(n_tasks) + " pools of stride " +
                                          std::to_string(pool_stride) + " need " + std::to_string(need) +
                                          " bytes of scratch, above the maximum " +
                                          std::to_string((size_t)FK_PRUNE_NOALLOC_SCRATCH_BYTES));
        if (p->flags & FK_FLAG_GRAM) {
            const uint32_t want = p->flags & (FK_FLAG_GRAM_BF16 | FK_FLAG_GRAM_F16);
            FK_CHECK_ARG(h->ds.gram_fmt != FK_GRAM_NONE && h->ds.gram_req == want,
                         "fk_prune_noalloc with FK_FLAG_GRAM needs the dataset's Gram form prepared by an earlier "
                         "fk_prune or fk_build with the same flags");
        }
        if (n_tasks == 0) {
            if (cc) *cc = 0;
            return;
        }
        select_device(h);
        FK_CHECK_ARG(h->noalloc != nullptr, "scratch address not set up");
        uint8_t *at = h->noalloc;
        auto take = [&](size_t bytes) {
            uint8_t *r = at;
            at += align256(bytes);
            return r;
        };
        const size_t n = n_tasks;
        auto *tally = reinterpret_cast<unsigned long long *>(take((size_t)FK_TALLY_WORDS * 8));
        auto *dpool = reinterpret_cast<uint64_t *>(take(n * pool_stride * 8));
        auto *dlen = reinterpret_cast<uint32_t *>(take(n * 4));
        auto *dq = reinterpret_cast<uint32_t *>(take(n * 4));
        auto *doid = reinterpret_cast<uint32_t *>(take(n * out_stride * 4));
        auto *dod = reinterpret_cast<float *>(take(n * out_stride * 4));
        auto *dolen = reinterpret_cast<uint32_t *>(take(n * 4));
        FK_HIP(hipMemcpyAsync(dpool, pools, n * pool_stride * 8, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(dlen, pool_len, n * 4, hipMemcpyHostToDevice, h->stream));
        if (q) FK_HIP(hipMemcpyAsync(dq, q, n * 4, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_fill_zero_u64, dim3(FK_TALLY_WORDS / 256), dim3(256), 0, h->stream, tally,
                           (uint32_t)FK_TALLY_WORDS);
        FK_LAUNCH_CHECK();
        PruneLaunch a = prune_args(h, type, pool_stride, n_tasks, oracle_size, p, reversed, out_stride);
        a.cand = dpool;
        a.cand_len = dlen;
        a.qid = q ? dq : nullptr;
        a.out_ids = doid;
        a.out_d = dod;
        a.out_len = dolen;
        a.cc = tally;
        launch_prune(h->stream, h->ds, a);
        FK_HIP(hipMemcpyAsync(out_ids, doid, n * out_stride * 4, hipMemcpyDeviceToHost, h->stream));
        if (out_dists) FK_HIP(hipMemcpyAsync(out_dists, dod, n * out_stride * 4, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipMemcpyAsync(out_len, dolen, n * 4, hipMemcpyDeviceToHost, h->stream));
        unsigned long long words[FK_TALLY_WORDS];
        FK_HIP(hipMemcpyAsync(words, tally, sizeof(words), hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipStreamSynchronize(h->stream));
        unsigned long long c = 0;
        for (uint32_t j = 0; j < FK_TALLY_STRIPES; ++j) c += words[j * FK_TALLY_STRIDE];
        if (cc) *cc = c;
    });
}

fk_status fk_knn_bruteforce(fk_handle *h, const float *queries, uint32_t n_q, uint32_t n_used, uint32_t K,
                            const uint32_t *self_ids, uint32_t *out_ids, float *out_dists) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        FK_CHECK_ARG(n_used >= 1 && n_used <= h->ds.n_total, "n_used out of range");
        FK_CHECK_ARG(K >= 1 && K < n_used + (self_ids ? 0u : 1u), "K must be below the number of rows");
        if (n_q == 0) return;
        FK_CHECK_ARG(queries && out_ids, "null pointer");
        select_device(h);
        const Dataset &ds = h->ds;
        DevBuf<float> dq((size_t)n_q * ds.dpad);
        upload_rows(h, dq.p, ds.dpad, queries, (size_t)ds.d * 4, n_q, ds.d);
        DevBuf<uint32_t> dself, dout((size_t)n_q * K);
        DevBuf<float> dd((size_t)n_q * K);
        if (self_ids) {
            dself.alloc(n_q);
            FK_HIP(hipMemcpyAsync(dself.p, self_ids, (size_t)n_q * 4, hipMemcpyHostToDevice, h->stream));
        }
        bruteforce_knn(h->stream, ds, dq.p, ds.dpad, n_q, n_used, K, self_ids ? dself.p : nullptr, dout.p, dd.p);
        FK_HIP(hipMemcpyAsync(out_ids, dout.p, (size_t)n_q * K * 4, hipMemcpyDeviceToHost, h->stream));
        if (out_dists)
            FK_HIP(hipMemcpyAsync(out_dists, dd.p, (size_t)n_q * K * 4, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipStreamSynchronize(h->stream));
    });
}

void fk_graph_free(fk_graph *g) {
    if (!g) return;
    free(g->offsets);
    free(g->ids);
    g->offsets = nullptr;
    g->ids = nullptr;
}

}  // extern "C"
