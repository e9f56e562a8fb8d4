// Occlusion prune with the occlusion distances in Gram form (prune variant
// 3, builds with FK_FLAG_GRAM on a prepared dataset, nsg_R <= 64, width <=
// 256).
//
// The decisions and distance counts are those of *_sync_prune /
// *_reversed_prune (src/kgraph.cpp:1809-1851 and the tau / alpha forms), as
// in k_prune: candidates are examined in pool order, each checked against
// the kept list in kept order until the first kept neighbour that has its
// id or occludes it.  One wavefront per pool takes the candidates 16 at a
// time and computes, before any decision, the Gram-form distance of every
// (candidate, kept neighbour) pair and of every pair within the batch as
// 16 x 16 MFMA tiles (fk_gram.h): on exact-form rows that value IS the
// canonical FP32 distance; on the bounded forms it brackets it (gram_lower /
// gram_upper), so a pair is decided by the bracket -- occluded, not occluded
// -- or, when the bracket straddles the rule's threshold, by the canonical
// distance computed in FP32 from the rows.  Each candidate's verdicts against
// the kept list are three 64-bit masks (occluded, undecided, same id); its
// check walks them in kept order, then the batch's accepts are resolved in
// candidate order with the in-batch verdicts, exactly as the sequential loop.
//
// Operands come from the dataset's 16-bit copy (rows16).  A batch's 16
// candidate rows are loaded once into registers as MFMA A fragments (lane l:
// row l & 15, values 32 s + 8 (l >> 4) .. +7 of K-step s) and serve as both
// operands of the in-batch tile; the kept tiles read their B fragments from
// rows16 (L2-resident for the pools of a workgroup's neighbourhood).  A cache
// of the kept rows in LDS was slower on every shape measured: R rows per
// wave leave two or three waves per SIMD where registers allow four.
#include "fk_gram.h"

namespace fk {

#define FK_PGR_WPB 4    // wavefronts (pools) per workgroup
#define FK_PGR_RMAX 64  // kept neighbours per pool (one 64-bit mask per candidate)
#define FK_PGR_NSMAX 8  // K-steps of 32 values: dpad <= 256
// 1 +- 2^-20 and 1 +- 2^-18 (exact in FP32): the margins of the square-root-
// free bracket tests (verdict in k_prune_gram)
#define FK_PGR_UP20 1.00000095367431640625f
#define FK_PGR_DN20 0.99999904632568359375f
#define FK_PGR_UP18 1.000003814697265625f
#define FK_PGR_DN18 0.999996185302734375f

// NS: K-steps of the tiles (ceil(dpad / 32), rounded up to 2, 3, 4 or 8;
// the values past dpad are zero).
template <int RULE, int FMT, int NS>
__global__ __launch_bounds__(64 * FK_PGR_WPB) __attribute__((amdgpu_waves_per_eu(NS <= 4 ? 4 : 3))) void k_prune_gram(const uint16_t *__restrict__ rows16,
                                                                const float *__restrict__ chains, uint32_t dpad,
                                                                const float2 *__restrict__ nrm, GramBound gb,
                                                                PruneLaunch a) {
    __shared__ uint32_t s_so[FK_PGR_WPB][FK_PGR_RMAX];
    __shared__ float s_sd[FK_PGR_WPB][FK_PGR_RMAX];
    __shared__ float2 s_sn[FK_PGR_WPB][FK_PGR_RMAX];
    __shared__ uint32_t s_cid[FK_PGR_WPB][16];
    __shared__ float s_cpd[FK_PGR_WPB][16];
    __shared__ float4 s_cb[FK_PGR_WPB][16];  // per candidate: T, U, |x~|^2, |x| (verdict)
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t vb = a.xcd_order ? xcd_position((gridDim.x + FK_NUM_XCD - 1) / FK_NUM_XCD) : blockIdx.x;
    const uint32_t task = vb * FK_PGR_WPB + w;
    if (task >= a.n_tasks) return;
    const uint32_t slot = a.task_list ? a.task_list[a.task_base + task] : a.task_base + task;
    const uint64_t *c = a.cand + (a.cand_off ? a.cand_off[slot] : (uint64_t)slot * a.cand_stride);
    const uint32_t len = a.cand_len[slot];
    const uint32_t q = a.qid ? a.qid[slot] : slot;
    uint32_t *oid = a.out_ids + (uint64_t)slot * a.out_stride;
    float *od = a.out_d ? a.out_d + (uint64_t)slot * a.out_stride : nullptr;
    uint32_t *so = s_so[w];
    float *sd = s_sd[w];
    float2 *sn = s_sn[w];
    uint32_t *cid = s_cid[w];
    float *cpd = s_cpd[w];
    float4 *cb = s_cb[w];
    const uint32_t R = a.R;
    const uint32_t nb = dpad >> 3, g = lane >> 3, k = lane & 7u;
    const uint32_t q8 = (lane >> 4) * 8u;
    // fragment s of this lane's quarter of a 16-bit row (zero past dpad)
    auto frag = [&](const uint16_t *r, uint32_t s) __attribute__((always_inline)) -> fk_s16x8 {
        const uint32_t kk = s * 32u + q8;
        if (kk < dpad) return *reinterpret_cast<const fk_s16x8 *>(r + kk);
        return fk_s16x8{0, 0, 0, 0, 0, 0, 0, 0};
    };

    uint32_t start = 0;
    if (len > 0 && key_id(c[0]) == q) start = 1;
    if (start >= len || R == 0) {
        if (lane == 0) a.out_len[slot] = 0;
        return;
    }
    const uint32_t limit = len < a.maxc ? len : a.maxc;
    {  // the first candidate is always kept
        const uint64_t k0 = c[start];
        const uint32_t id0 = key_id(k0);  // kept whatever its id, with its row (as k_prune)
        if (lane == 0) {
            so[0] = id0;
            sd[0] = key_dist(k0);
            sn[0] = nrm[id0];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t nout = 1;
    bool stop = nout >= R;
    unsigned long long cc = 0, exam = 0, unc = 0;
    // the verdict of one (checked candidate, kept) pair from its Gram-form
    // distance dt: 1 occluded, 0 not, -1 undecided (bounded forms only).
    // cbi = the checked candidate's (T, U, |x~|^2, |x|), T and U the square
    // roots of its rule threshold th (pdc, or pdc - 3 tau) over (1 - gam)
    // and (1 + gam).  The bracket's tests gram_lower(dt) >= th and
    // gram_upper(dt) < th are taken without a square root per pair: with
    // f, e the error terms of gram_lower, sqrt(max(dt - f, 0)) - e >= T iff
    // dt - f >= (e + T)^2, and sqrt(dt + f) + e < U iff U > e and
    // dt + f < (U - e)^2.  T is rounded up and U down by 2^-20 when made,
    // e up by 2^-20 here, and each compared side moved by 2^-18, more than
    // the rounding of these few FP32 operations, so either test implies the
    // original one; th <= 0 cannot be met by a distance (not occluded).
    auto verdict = [&](float dt, float4 cbi, float2 kn, float pdc, float dqt) __attribute__((always_inline)) -> int {
        if (gb.exact) return occludes<RULE>(dt, pdc, dqt, a.tau3, a.cos_thr) ? 1 : 0;
        const float th = RULE == FK_RULE_TAU ? pdc - a.tau3 : pdc;
        const float f = gb.f_rel * (cbi.z + kn.x);
        const float e = (gb.u * (cbi.w + kn.y) + gb.e_abs) * FK_PGR_UP20;
        const float lo = cbi.x + e;
        const bool not_occl = !(th > 0.f) || (dt - f) * FK_PGR_DN18 >= lo * lo * FK_PGR_UP18;
        bool occl = false;
        if (RULE != FK_RULE_ALPHA) {
            const float hi = cbi.y - e;
            occl = hi > 0.f && (dt + f) * FK_PGR_UP18 < hi * hi * FK_PGR_DN18;
        }
        return not_occl ? 0 : occl ? 1 : -1;
    };
    // canonical FP32 distance of two rows (chain-major copies), one octet
    auto exact = [&](uint32_t x, uint32_t y) __attribute__((always_inline)) -> float {
        return octet_dist2<0>(chains + (size_t)x * dpad + k * nb, chains + (size_t)y * dpad + k * nb, nb);
    };
    // The 16-bit fields of this octet's candidates g and g + 8 (rows of the
    // tile) in the ballots of a tile (lane l holds row 4 (l >> 4) + e, column
    // l & 15): both in ballot g & 3, at bits 16 (g >> 2) and 16 (g >> 2) + 32.
    // The ballot is picked with masks, not a branch per octet.
    const uint64_t sel1 = 0ull - (uint64_t)(g & 1u), sel2 = 0ull - (uint64_t)((g >> 1) & 1u);
    const uint32_t fsh = 16u * (g >> 2);
    auto fields = [&](const uint64_t (&bl)[4], uint64_t &f0, uint64_t &f1) __attribute__((always_inline)) {
        const uint64_t b01 = bl[0] ^ ((bl[0] ^ bl[1]) & sel1);
        const uint64_t b23 = bl[2] ^ ((bl[2] ^ bl[3]) & sel1);
        const uint64_t b = (b01 ^ ((b01 ^ b23) & sel2)) >> fsh;
        f0 = b & 0xFFFFull;
        f1 = (b >> 32) & 0xFFFFull;
    };

    // Batches of 16 candidates, software-pipelined: the keys of batch b + 1
    // are loaded when batch b starts, their norms after its kept tiles and,
    // for rows of up to 3 K-steps (PFA; not the tau rule), their A fragments
    // after its in-batch tile (af is free from there on; otherwise they are
    // loaded when the batch starts, which keeps the kernel within four waves
    // per SIMD without spills).
    // Lane l holds candidate ci = l & 15 (row 0 for an empty or out-of-range
    // slot; its verdicts are never read).
    const uint32_t ci = lane & 15u;
    auto key_at = [&](uint32_t p) __attribute__((always_inline)) -> uint64_t {
        return p < limit ? c[p] : FK_KEY_EMPTY;
    };
    auto row_id = [&](uint64_t key) __attribute__((always_inline)) -> uint32_t {
        const uint32_t id = key_id(key);
        return id < a.oracle_n ? id : 0u;
    };
    fk_s16x8 af[NS];
    auto load_a = [&](uint64_t key) __attribute__((always_inline)) {
        const uint16_t *ar = rows16 + (size_t)row_id(key) * dpad;
#pragma unroll
        for (uint32_t s = 0; s < NS; ++s) af[s] = frag(ar, s);
    };
    constexpr bool PFA = NS <= 3 && RULE != FK_RULE_TAU;
    uint64_t key_n = key_at(start + 1 + ci);
    float2 nrm_n = nrm[row_id(key_n)];
    if (PFA) load_a(key_n);
    for (uint32_t pos = start + 1; !stop && pos < limit; pos += 16) {
        const uint64_t key = key_n;
        const float2 n = nrm_n;
        if (!PFA) load_a(key);
        key_n = key_at(pos + 16u + ci);
        {
            const uint32_t pid = key_id(key);
            const bool valid = pid < a.oracle_n;
            if (lane < 16) {
                const float pd = key_dist(key);
                const float th = RULE == FK_RULE_TAU ? pd - a.tau3 : pd;
                const float tp = th > 0.f ? th : 0.f;
                cid[ci] = valid ? pid : 0xFFFFFFFFu;
                cpd[ci] = pd;
                cb[ci] = make_float4(__builtin_sqrtf(tp / (1.f - gb.gam)) * FK_PGR_UP20,
                                     __builtin_sqrtf(tp / (1.f + gb.gam)) * FK_PGR_DN20, n.x, n.y);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint32_t n0 = nout;
        // verdict masks of this octet's two candidates g and g + 8
        uint64_t m_occ[2] = {0, 0}, m_und[2] = {0, 0}, m_dup[2] = {0, 0};
        // B fragments of kept tile kt (row kt * 16 + ci of the kept list, row
        // 0 past its end); for narrow rows the next tile's are loaded while
        // this one's verdicts are made (where the registers for them would
        // not cost a wave per SIMD)
        constexpr bool PF = NS <= 3 || (NS == 4 && RULE != FK_RULE_ALPHA);
        fk_s16x8 bf[NS];
        auto load_b = [&](uint32_t kt) __attribute__((always_inline)) {
            const uint32_t kj = kt * 16u + ci;
            const uint16_t *br = rows16 + (size_t)so[kj < n0 ? kj : 0u] * dpad;
#pragma unroll
            for (uint32_t s = 0; s < NS; ++s) bf[s] = frag(br, s);
        };
        if (PF) load_b(0);
        for (uint32_t kt = 0; kt * 16u < n0; ++kt) {
            fk_f32x4 gt = {0.f, 0.f, 0.f, 0.f};
            if (!PF) load_b(kt);
#pragma unroll
            for (uint32_t s = 0; s < NS; ++s) gt = mfma_step16<FMT>(af[s], bf[s], gt);
            if (PF && (kt + 1u) * 16u < n0) load_b(kt + 1u);
            const uint32_t j = kt * 16u + (lane & 15u);
            const bool jok = j < n0;
            const uint32_t jj = jok ? j : 0u;
            const uint32_t jid = so[jj];
            const float dqt = sd[jj];
            const float2 kn = sn[jj];
            uint64_t bo[4], bu[4], bd[4];
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                const uint32_t i = 4u * (lane >> 4) + e;
                const uint32_t iid = cid[i];
                const float4 cbi = cb[i];
                const float dt = (cbi.z + kn.x) - 2.f * (gt[e] * gb.inv_s2);
                const int v = verdict(dt, cbi, kn, cpd[i], dqt);
                const bool ok = jok && iid != 0xFFFFFFFFu;
                const bool d = ok && jid == iid;
                const bool o = ok && !d && v == 1;
                const bool u = ok && !d && v < 0;
                bo[e] = wave_ballot(o);
                bu[e] = wave_ballot(u);
                bd[e] = wave_ballot(d);
            }
            uint64_t f0, f1;
            fields(bo, f0, f1);
            m_occ[0] |= f0 << (16u * kt);
            m_occ[1] |= f1 << (16u * kt);
            fields(bu, f0, f1);
            m_und[0] |= f0 << (16u * kt);
            m_und[1] |= f1 << (16u * kt);
            fields(bd, f0, f1);
            m_dup[0] |= f0 << (16u * kt);
            m_dup[1] |= f1 << (16u * kt);
        }
        nrm_n = nrm[row_id(key_n)];
        // in-batch verdicts: checker r (tile row) against candidate c (column)
        uint64_t bo2[4], bu2[4], bd2[4];
        {
            fk_f32x4 gt = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (uint32_t s = 0; s < NS; ++s) gt = mfma_step16<FMT>(af[s], af[s], gt);
            if (PFA) load_a(key_n);
            const uint32_t cc_ = lane & 15u;
            const uint32_t cidc = cid[cc_];
            const float pdc_ = cpd[cc_];
            const float4 cbc = cb[cc_];
            const float2 ncc = make_float2(cbc.z, cbc.w);
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                const uint32_t r = 4u * (lane >> 4) + e;
                const uint32_t rid = cid[r];
                const float4 cbr = cb[r];
                const float dt = (cbr.z + ncc.x) - 2.f * (gt[e] * gb.inv_s2);
                const int v = verdict(dt, cbr, ncc, cpd[r], pdc_);
                const bool ok = r > cc_ && rid != 0xFFFFFFFFu && cidc != 0xFFFFFFFFu;
                const bool d = ok && rid == cidc;
                const bool o = ok && !d && v == 1;
                const bool u = ok && !d && v < 0;
                bo2[e] = wave_ballot(o);
                bu2[e] = wave_ballot(u);
                bd2[e] = wave_ballot(d);
            }
        }
        // each octet walks its two candidates' masks in kept order
        bool occl[2];
        uint32_t cnt[2], ncu[2];
        const uint64_t lowm = n0 >= 64u ? ~0ull : ((1ull << n0) - 1ull);
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h) {
            const uint32_t i = g + 8u * h;
            const uint32_t iid = cid[i];
            occl[h] = false;
            cnt[h] = 0;
            ncu[h] = 0;
            if (iid == 0xFFFFFFFFu) continue;  // octet-uniform
            uint64_t m = (m_occ[h] | m_und[h] | m_dup[h]) & lowm;
            uint32_t stop_t = n0;
            while (m) {
                const uint32_t t = (uint32_t)__builtin_ctzll(m);
                const uint64_t bit = 1ull << t;
                if (m_dup[h] & bit) { occl[h] = true; stop_t = t; break; }
                if (m_occ[h] & bit) { occl[h] = true; stop_t = t + 1; break; }
                // undecided: the canonical distance
                ++ncu[h];
                if (occludes<RULE>(exact(iid, so[t]), cpd[i], sd[t], a.tau3, a.cos_thr)) {
                    occl[h] = true;
                    stop_t = t + 1;
                    break;
                }
                m &= m - 1;
            }
            cnt[h] = stop_t;  // kept neighbours checked (a same-id stop counts no distance)
        }
        // accepts in candidate order; later open candidates are checked
        // against each newly kept one (its in-batch verdict)
        uint64_t i_dup[2], i_occ[2], i_und[2];  // in-batch verdicts of candidates g, g + 8
        fields(bd2, i_dup[0], i_dup[1]);
        fields(bo2, i_occ[0], i_occ[1]);
        fields(bu2, i_und[0], i_und[1]);
        uint32_t done = 0, stop_i = 15;
        while (true) {
            const uint64_t ob = wave_ballot(k == 0 && cid[g] != 0xFFFFFFFFu && !occl[0]) |
                                (wave_ballot(k == 0 && cid[g + 8] != 0xFFFFFFFFu && !occl[1]) << 1);
            // candidate i is open when bit 8 * (i & 7) + (i >> 3) is set
            uint32_t ai = 0xFFFFFFFFu;
            for (uint32_t i = done; i < 16; ++i)
                if ((ob >> (8u * (i & 7u) + (i >> 3))) & 1ull) { ai = i; break; }
            if (ai == 0xFFFFFFFFu) break;
            const uint32_t slot_o = nout;
            if (lane == 0) {
                so[slot_o] = cid[ai];
                sd[slot_o] = cpd[ai];
                sn[slot_o] = make_float2(cb[ai].z, cb[ai].w);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            ++nout;
            done = ai + 1;
            if (nout >= R) { stop = true; stop_i = ai; break; }
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                const uint32_t r = g + 8u * h;
                if (r > ai && cid[r] != 0xFFFFFFFFu && !occl[h]) {  // octet-uniform
                    const uint64_t bit = 1ull << ai;
                    if (i_dup[h] & bit) {
                        occl[h] = true;
                    } else {
                        ++cnt[h];
                        if (i_occ[h] & bit) {
                            occl[h] = true;
                        } else if (i_und[h] & bit) {
                            ++ncu[h];
                            if (occludes<RULE>(exact(cid[r], cid[ai]), cpd[r], cpd[ai], a.tau3, a.cos_thr))
                                occl[h] = true;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h) {
            const uint32_t i = g + 8u * h;
            if (k == 0 && i <= stop_i && cid[i] != 0xFFFFFFFFu) {
                cc += cnt[h];
                exam += 1;
                // bounded forms: every counted distance not computed in FP32
                if (!gb.exact) unc += cnt[h] - ncu[h];
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_wave_barrier();
    for (uint32_t t = lane; t < nout; t += 64) {
        oid[t] = so[t];
        if (od) od[t] = sd[t];
    }
    if (lane == 0) a.out_len[slot] = nout;
    if (a.cc) {
        const unsigned long long s = wave_sum(cc);
        if (lane == 0 && s) tally_add(a.cc, s);
    }
    if (a.examined) {
        const unsigned long long s = wave_sum(exam);
        if (lane == 0 && s) tally_add(a.examined, s);
    }
    if (a.uncomputed) {
        const unsigned long long s = wave_sum(unc);
        if (lane == 0 && s) tally_add(a.uncomputed, s);
    }
}

// K-steps the kernel is instantiated for, for rows of dpad values (0: none).
static uint32_t pgr_steps(uint32_t dpad) {
    const uint32_t ns = (dpad + 31u) / 32u;
    return ns <= 2 ? 2u : ns <= 4 ? ns : ns <= FK_PGR_NSMAX ? 8u : 0u;
}

// Variant 3 applies to a.gram launches on a dataset with its Gram form and
// 16-bit copy prepared, nsg_R <= 64 and dpad <= 256.
bool prune_gram_applies(const Dataset &ds, const PruneLaunch &a) {
    return a.gram && ds.gram_fmt != FK_GRAM_NONE && a.R <= FK_PGR_RMAX && ds.gram_nrm.p && ds.rows16.p &&
           pgr_steps(ds.dpad) != 0;
}

template <int FMT, int NS>
static void launch_pgr(hipStream_t s, const Dataset &ds, const PruneLaunch &a, const GramBound &gb, unsigned grid) {
    switch (a.rule) {
        case FK_RULE_NSG:
            hipLaunchKernelGGL((k_prune_gram<FK_RULE_NSG, FMT, NS>), dim3(grid), dim3(64 * FK_PGR_WPB), 0, s,
                               ds.rows16.p, ds.chains.p, ds.dpad, ds.gram_nrm.p, gb, a);
            break;
        case FK_RULE_TAU:
            hipLaunchKernelGGL((k_prune_gram<FK_RULE_TAU, FMT, NS>), dim3(grid), dim3(64 * FK_PGR_WPB), 0, s,
                               ds.rows16.p, ds.chains.p, ds.dpad, ds.gram_nrm.p, gb, a);
            break;
        default:
            hipLaunchKernelGGL((k_prune_gram<FK_RULE_ALPHA, FMT, NS>), dim3(grid), dim3(64 * FK_PGR_WPB), 0, s,
                               ds.rows16.p, ds.chains.p, ds.dpad, ds.gram_nrm.p, gb, a);
            break;
    }
}

template <int FMT>
static void launch_pgr_fmt(hipStream_t s, const Dataset &ds, const PruneLaunch &a, const GramBound &gb,
                           unsigned grid) {
    switch (pgr_steps(ds.dpad)) {
        case 2: launch_pgr<FMT, 2>(s, ds, a, gb, grid); break;
        case 3: launch_pgr<FMT, 3>(s, ds, a, gb, grid); break;
        case 4: launch_pgr<FMT, 4>(s, ds, a, gb, grid); break;
        default: launch_pgr<FMT, 8>(s, ds, a, gb, grid); break;
    }
}

// Launches variant 3 when it applies; returns false, launching nothing,
// otherwise.
bool launch_prune_gram(hipStream_t s, const Dataset &ds, const PruneLaunch &a) {
    if (!prune_gram_applies(ds, a)) return false;
    const GramBound gb = gram_bound(ds);
    unsigned grid = (a.n_tasks + FK_PGR_WPB - 1) / FK_PGR_WPB;
    if (a.xcd_order) grid = (grid + FK_NUM_XCD - 1) / FK_NUM_XCD * FK_NUM_XCD;
    if (ds.gram_fmt == FK_GRAM_F16) launch_pgr_fmt<FK_GRAM_F16>(s, ds, a, gb, grid);
    else launch_pgr_fmt<FK_GRAM_BF16>(s, ds, a, gb, grid);
    FK_LAUNCH_CHECK();
    return true;
}

}  // namespace fk
