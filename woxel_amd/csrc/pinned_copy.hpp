// Large copies between pageable host memory and device memory through two pinned
// buffers: the DMA of one piece runs while host threads copy the next (or previous)
// piece between pageable memory and the other pinned buffer. A plain hipMemcpy of
// pageable memory stages through the runtime's buffers on one thread; writing fresh
// host memory (first touch, page faults) on one thread is its slow part.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace wx {

class PinnedCopier {
  public:
    static constexpr size_t kPiece = 64ull << 20;
    // Copies below this size take a plain hipMemcpy (the pinned buffers are not worth it).
    static constexpr size_t kMin = 32ull << 20;

    // Copies are ordered on `stream` (0: a stream of the copier's own).
    explicit PinnedCopier(hipStream_t stream = nullptr) {
        const unsigned hc = std::thread::hardware_concurrency();
        threads_ = (int)std::min(16u, std::max(1u, hc));
        if (!stream) {
            ok_ = hipStreamCreateWithFlags(&own_, hipStreamNonBlocking) == hipSuccess;
            st_ = own_;
        } else {
            st_ = stream;
        }
        for (int i = 0; i < 2 && ok_; ++i)
            ok_ = hipHostMalloc(reinterpret_cast<void **>(&buf_[i]), kPiece, hipHostMallocDefault) == hipSuccess &&
                  hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming) == hipSuccess;
    }
    ~PinnedCopier() {
        if (st_) (void)hipStreamSynchronize(st_);
        for (int i = 0; i < 2; ++i) {
            if (buf_[i]) (void)hipHostFree(buf_[i]);
            if (ev_[i]) (void)hipEventDestroy(ev_[i]);
        }
        if (own_) (void)hipStreamDestroy(own_);
    }
    PinnedCopier(const PinnedCopier &) = delete;
    PinnedCopier &operator=(const PinnedCopier &) = delete;
    bool ok() const { return ok_; }

    // dst (host) <- src (device); complete when it returns true.
    bool to_host(void *dst, const void *src, size_t bytes) {
        size_t prev_off = 0, prev_len = 0;
        int prev = -1;
        for (size_t off = 0, i = 0; off < bytes; off += kPiece, ++i) {
            const int b = (int)(i & 1u);
            const size_t len = std::min(kPiece, bytes - off);
            // buffer b was last read by the host copy two pieces back, already done
            if (hipMemcpyAsync(buf_[b], static_cast<const char *>(src) + off, len, hipMemcpyDeviceToHost, st_) !=
                    hipSuccess ||
                hipEventRecord(ev_[b], st_) != hipSuccess)
                return false;
            if (prev >= 0) {
                if (hipEventSynchronize(ev_[prev]) != hipSuccess) return false;
                spread(static_cast<char *>(dst) + prev_off, buf_[prev], prev_len);
            }
            prev = b, prev_off = off, prev_len = len;
        }
        if (prev < 0) return true;
        if (hipEventSynchronize(ev_[prev]) != hipSuccess) return false;
        spread(static_cast<char *>(dst) + prev_off, buf_[prev], prev_len);
        return true;
    }

    // dst (device) <- src (host); complete when it returns true.
    bool to_device(void *dst, const void *src, size_t bytes) {
        bool used[2] = {false, false};
        for (size_t off = 0, i = 0; off < bytes; off += kPiece, ++i) {
            const int b = (int)(i & 1u);
            const size_t len = std::min(kPiece, bytes - off);
            // buffer b is free once the DMA two pieces back has read it
            if (used[b] && hipEventSynchronize(ev_[b]) != hipSuccess) return false;
            spread(buf_[b], static_cast<const char *>(src) + off, len);
            if (hipMemcpyAsync(static_cast<char *>(dst) + off, buf_[b], len, hipMemcpyHostToDevice, st_) !=
                    hipSuccess ||
                hipEventRecord(ev_[b], st_) != hipSuccess)
                return false;
            used[b] = true;
        }
        return hipStreamSynchronize(st_) == hipSuccess;
    }

  private:
    // memcpy split over the copier's threads (the calling thread takes the first part)
    void spread(char *dst, const char *src, size_t len) {
        const size_t T = std::min<size_t>((size_t)threads_, len / (1u << 20) + 1);
        std::vector<std::thread> th;
        for (size_t t = 1; t < T; ++t)
            th.emplace_back([=] { std::memcpy(dst + len * t / T, src + len * t / T, len * (t + 1) / T - len * t / T); });
        std::memcpy(dst, src, len / T);
        for (auto &x : th) x.join();
    }

    hipStream_t st_ = nullptr, own_ = nullptr;
    int threads_ = 1;
    bool ok_ = true;
    char *buf_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
};

}  // namespace wx
