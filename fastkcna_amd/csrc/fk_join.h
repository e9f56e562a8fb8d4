// The local join's distance block, shared by the NN-descent join and the
// fk_join_distances entry point.
//
// src/kgraph.cpp:970-987: for one node, every pair (x, y) of its sampled
// list with x among the first nn ("new") entries and y > x.  The rows are
// staged in LDS (row stride dpad + 4 floats, so 16-byte reads of
// consecutive rows fall in different banks) and one thread computes one
// pair.  A list too long for the staging area is done in tiles: a chunk of
// x rows stays resident while chunks of y rows stream through the rest.
#pragma once
#include "fk_device.h"

#define FK_JOIN_LDS_ROWS_BYTES (40 * 1024)
#ifndef FK_JOIN_SMALL_ROWS
#define FK_JOIN_SMALL_ROWS 40  // lists this short are joined with a small staging area
#endif

__device__ __forceinline__ void join_stage_rows(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                const uint32_t *s_id, uint32_t first, uint32_t count,
                                                float *s_dst) {
    const uint32_t ls = dpad + 4, v4 = dpad >> 2;
    for (uint32_t i = threadIdx.x; i < count * v4; i += blockDim.x) {
        const uint32_t r = i / v4, c = i - r * v4;
        *reinterpret_cast<float4 *>(s_dst + r * ls + c * 4) =
            *reinterpret_cast<const float4 *>(rows + (size_t)s_id[first + r] * stride + c * 4);
    }
}

// Calls emit(x, y, dist) once for every pair of the block.  s_id holds the
// T list ids in LDS; s_rows has rows_bytes (at most FK_JOIN_LDS_ROWS_BYTES).
// Every thread of the block must call it; it returns after a barrier-free
// compute phase, so the caller synchronises before reading what emit wrote
// to LDS.
template <class Emit>
__device__ void join_block(const float *__restrict__ rows, uint32_t stride, uint32_t dpad, const uint32_t *s_id,
                           uint32_t T, uint32_t nn, float *s_rows, Emit emit,
                           uint32_t rows_bytes = FK_JOIN_LDS_ROWS_BYTES) {
    const uint32_t ls = dpad + 4;
    const uint32_t cap = rows_bytes / (ls * 4);
    if (T <= cap) {
        join_stage_rows(rows, stride, dpad, s_id, 0, T, s_rows);
        __syncthreads();
        const uint32_t cells = nn * T;
        for (uint32_t c = threadIdx.x; c < cells; c += blockDim.x) {
            const uint32_t x = c / T, y = c - x * T;
            if (y <= x) continue;
            emit(x, y, l2_canon(s_rows + x * ls, s_rows + y * ls, dpad));
        }
        return;
    }
    const uint32_t bx = nn < cap / 2 ? nn : cap / 2, by = cap - bx;
    float *s_y = s_rows + bx * ls;
    for (uint32_t x0 = 0; x0 < nn; x0 += bx) {
        const uint32_t nx = (nn - x0) < bx ? (nn - x0) : bx;
        __syncthreads();  // the previous tiles are done with the x rows
        join_stage_rows(rows, stride, dpad, s_id, x0, nx, s_rows);
        for (uint32_t y0 = x0 + 1; y0 < T; y0 += by) {
            const uint32_t ny = (T - y0) < by ? (T - y0) : by;
            if (y0 != x0 + 1) __syncthreads();  // the previous tile is done with the y rows
            join_stage_rows(rows, stride, dpad, s_id, y0, ny, s_y);
            __syncthreads();
            const uint32_t cells = nx * ny;
            for (uint32_t c = threadIdx.x; c < cells; c += blockDim.x) {
                const uint32_t xi = c / ny, yi = c - xi * ny;
                const uint32_t x = x0 + xi, y = y0 + yi;
                if (y <= x) continue;
                emit(x, y, l2_canon(s_rows + xi * ls, s_y + yi * ls, dpad));
            }
        }
    }
}

// Position of pair (x, y) in a node's pair array: rows x ascending, each
// holding y = x + 1 .. T - 1.
__device__ __forceinline__ uint64_t join_pair_index(uint32_t x, uint32_t y, uint32_t T) {
    return (uint64_t)x * (T - 1) - (uint64_t)x * (x - 1) / 2 + (y - x - 1);
}
