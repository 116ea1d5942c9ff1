// Host/device buffers of the drop-in API, the HIP counterparts of the
// reference's smart_cpu_buffer / smart_gpu_buffer / smart_gpu_cpu_buffer
// (CUDA_memory.h:49-210).  Like the reference they are shallow handles with
// an explicit destroy(); unlike it, a failed allocation throws
// std::runtime_error instead of leaving a half-built object.
#pragma once
#include <stddef.h>
#include <string.h>

#include <stdexcept>
#include <string>

#include "../adaptivemesh.h"

namespace am {

inline void check(int rc, const char *what)
{
    if (rc != AM_OK) throw std::runtime_error(std::string(what) + ": " + am_last_error_string());
}

template <class T>
struct smart_cpu_buffer {
    size_t dedicated_len = 0;
    T *cpu_buffer_ptr = nullptr;
    bool created = false;

    smart_cpu_buffer() = default;
    explicit smart_cpu_buffer(size_t n) : dedicated_len(n), cpu_buffer_ptr(new T[n]), created(true) {}
    size_t total_size() const { return dedicated_len * sizeof(T); }
    T &operator[](size_t i) { return cpu_buffer_ptr[i]; }
    const T &operator[](size_t i) const { return cpu_buffer_ptr[i]; }
    void destroy()
    {
        if (created) delete[] cpu_buffer_ptr;
        cpu_buffer_ptr = nullptr;
        created = false;
    }
};

template <class T>
struct smart_gpu_buffer {
    size_t dedicated_len = 0;
    T *gpu_buffer_ptr = nullptr;
    bool created = false;

    smart_gpu_buffer() = default;
    explicit smart_gpu_buffer(size_t n) : dedicated_len(n)
    {
        void *p = nullptr;
        check(am_device_alloc(&p, n * sizeof(T) + 16), "smart_gpu_buffer");
        gpu_buffer_ptr = static_cast<T *>(p);
        created = true;
    }
    size_t total_size() const { return dedicated_len * sizeof(T); }
    void swap_pointers(smart_gpu_buffer<T> &o)
    {
        if (dedicated_len != o.dedicated_len) throw std::runtime_error("cannot swap buffers of differing lengths");
        T *t = gpu_buffer_ptr;
        gpu_buffer_ptr = o.gpu_buffer_ptr;
        o.gpu_buffer_ptr = t;
    }
    void zero(void *stream = nullptr) { check(am_memset_async(gpu_buffer_ptr, 0, total_size(), stream), "zero"); }
    void destroy()
    {
        if (created) am_device_free(gpu_buffer_ptr);
        gpu_buffer_ptr = nullptr;
        created = false;
    }
};

// Host array plus a device mirror.  The device side is allocated on the
// first upload, so the octree logic runs (and is tested) without a GPU.
template <class T>
struct smart_gpu_cpu_buffer {
    size_t dedicated_len = 0;
    T *cpu_buffer_ptr = nullptr;
    T *gpu_buffer_ptr = nullptr;

    smart_gpu_cpu_buffer() = default;
    explicit smart_gpu_cpu_buffer(size_t n) : dedicated_len(n), cpu_buffer_ptr(new T[n]) {}
    size_t total_size() const { return dedicated_len * sizeof(T); }
    void ensure_gpu()
    {
        if (gpu_buffer_ptr) return;
        void *p = nullptr;
        check(am_device_alloc(&p, total_size() + 16), "smart_gpu_cpu_buffer");
        gpu_buffer_ptr = static_cast<T *>(p);
    }
    void copy_to_gpu(void *stream = nullptr)
    {
        ensure_gpu();
        check(am_memcpy_h2d(gpu_buffer_ptr, cpu_buffer_ptr, total_size(), stream), "copy_to_gpu");
        check(am_stream_sync(stream), "copy_to_gpu");  // the host copy may change right after
    }
    void copy_to_cpu(void *stream = nullptr)
    {
        check(am_memcpy_d2h(cpu_buffer_ptr, gpu_buffer_ptr, total_size(), stream), "copy_to_cpu");
    }
    void destroy()
    {
        delete[] cpu_buffer_ptr;
        cpu_buffer_ptr = nullptr;
        if (gpu_buffer_ptr) am_device_free(gpu_buffer_ptr);
        gpu_buffer_ptr = nullptr;
    }
};

// Device-only array allocated on first use (so the octree logic that owns
// one still runs without a GPU), grown when a larger length is asked for.
template <class T>
struct lazy_gpu_buffer {
    size_t dedicated_len = 0;
    T *gpu_buffer_ptr = nullptr;
    T *ensure(size_t n)
    {
        if (gpu_buffer_ptr && n <= dedicated_len) return gpu_buffer_ptr;
        if (gpu_buffer_ptr) am_device_free(gpu_buffer_ptr);
        gpu_buffer_ptr = nullptr;
        void *p = nullptr;
        check(am_device_alloc(&p, n * sizeof(T) + 16), "lazy_gpu_buffer");
        gpu_buffer_ptr = static_cast<T *>(p);
        dedicated_len = n;
        return gpu_buffer_ptr;
    }
    void destroy()
    {
        if (gpu_buffer_ptr) am_device_free(gpu_buffer_ptr);
        gpu_buffer_ptr = nullptr;
        dedicated_len = 0;
    }
};

// cuda_sync (CUDA_memory.h:213-220): wait for every launch on the device
inline int cuda_sync() { return am_device_sync(); }

// round_robin_threads (AMR.h:852-878): eight streams handed out in turn, so
// code written for the reference's plugins passes them through unchanged.
// The wrappers here take a stream as void*.
struct round_robin_threads {
    void *streams[8] = {};
    uint32_t current_stream_idx = 0;
    round_robin_threads()
    {
        for (int i = 0; i < 8; i++) check(am_stream_create(&streams[i]), "round_robin_threads");
    }
    round_robin_threads(const round_robin_threads &) = delete;
    round_robin_threads &operator=(const round_robin_threads &) = delete;
    ~round_robin_threads()
    {
        for (int i = 0; i < 8; i++) am_stream_destroy(streams[i]);
    }
    void set_stream_idx(uint32_t idx) { current_stream_idx = idx & 7u; }
    void *current_stream() const { return streams[current_stream_idx & 7u]; }
    void *next_stream() const { return streams[(current_stream_idx + 1u) & 7u]; }
    void *yield_stream()
    {
        current_stream_idx &= 7u;
        return streams[current_stream_idx++];
    }
};

}  // namespace am
