// Shared helpers of the C-ABI translation units.
#pragma once
#include <new>

#include "fk_internal.h"

namespace fk {

// Makes the handle's memory cache the one device buffers use while a call
// runs on this thread.
struct CacheScope {
    MemCache *prev;
    explicit CacheScope(MemCache *c) : prev(current_cache()) { current_cache() = c; }
    ~CacheScope() { current_cache() = prev; }
};

// Runs f, converting any failure into a status and the handle's last error.
template <class F>
fk_status guard(fk_handle *h, F &&f) {
    CacheScope scope(h ? &h->cache : nullptr);
    try {
        f();
        if (h) h->last_error.clear();
        return FK_OK;
    } catch (const Error &e) {
        if (h) h->last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        if (h) h->last_error = "host allocation failed";
        return FK_E_NOMEM;
    } catch (const std::exception &e) {
        if (h) h->last_error = e.what();
        return FK_E_INTERNAL;
    }
}

float alpha_cos_threshold(float angle);
int rule_of(fk_pg_type t);
void select_device(fk_handle *h);
void require_dataset(fk_handle *h);
void download_bytes(fk_handle *h, void *dst, const void *src, size_t bytes);
void *host_result_alloc(size_t bytes);
void upload_rows(fk_handle *h, float *dst, uint32_t dpad, const float *src, size_t src_stride_bytes, uint32_t n,
                 uint32_t d, float *colsum = nullptr);
// fk_prune_noalloc's scratch: its device address on `device` (current), and
// the bytes a call needs.
uint8_t *noalloc_scratch(int device);
size_t noalloc_bytes(uint32_t n_tasks, uint32_t pool_stride, uint32_t out_stride);

}  // namespace fk
