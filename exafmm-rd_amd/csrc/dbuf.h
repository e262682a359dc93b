// Growable device allocation owned by a plan.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace fmm3dpp {

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { if (p) (void)hipFree(p); }
    // Ensure capacity; contents are not preserved.
    hipError_t ensure(size_t b) {
        if (b <= bytes && p) return hipSuccess;
        if (p) {   // work queued on any stream may still read the old block
            (void)hipDeviceSynchronize();
            (void)hipFree(p);
            p = nullptr;
            bytes = 0;
        }
        size_t want = std::max<size_t>(b, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    // Ensure capacity keeping the first `keep` bytes (stream-ordered copy).
    hipError_t grow_keep(size_t b, size_t keep, hipStream_t st) {
        if (b <= bytes && p) return hipSuccess;
        size_t want = std::max<size_t>(std::max(b, bytes + bytes / 2), 256);
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, want);
        if (e != hipSuccess) return e;
        if (p && keep) {
            e = hipMemcpyAsync(q, p, keep, hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) { (void)hipFree(q); return e; }
            e = hipStreamSynchronize(st);
            if (e != hipSuccess) { (void)hipFree(q); return e; }
        }
        if (p) (void)hipFree(p);
        p = q;
        bytes = want;
        return hipSuccess;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Copies a host vector into d, ordered on st, and returns after the copy has completed: callers
// pass temporaries that are destroyed right after the call, and a pageable-memory copy may read
// its source only when the stream reaches it.
template <class T>
hipError_t upload(DBuf& d, const std::vector<T>& h, hipStream_t st) {
    hipError_t e = d.ensure(h.size() * sizeof(T));
    if (e != hipSuccess || h.empty()) return e;
    e = hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

}  // namespace fmm3dpp
