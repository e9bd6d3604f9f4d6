/*
 * runtime.cpp -- errors, devices, memory and events of libsygraph_hip.
 *
 * Replaces the SYCL runtime services the reference leans on: sycl::queue
 * construction and device queries (utils/device.hpp:16-34), USM allocation
 * (utils/memory.hpp:12-48), queue.copy/fill, and event timing
 * (sycl/event.hpp:19-44).
 */
#include <cstring>

#include "internal.hpp"

namespace {
thread_local std::string g_last_error = "no error";

/* Plain streaming copy, 16 bytes per lane per step: the reference point for
 * HBM bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE counters. */
__global__ void __launch_bounds__(256) copy_kernel_u128(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

/* Random 4-byte gather, dst[i] = src[h(i) mod src_words] with a splitmix64
 * hash of (i, seed): the reference point for counters read on random probes,
 * where each 4-byte read brings in a whole line.  Every index lies in
 * [0, src_words). */
__global__ void __launch_bounds__(256) gather_kernel_u32(const uint32_t* __restrict__ src, uint64_t src_words,
                                                         uint32_t* __restrict__ dst, size_t n, uint64_t seed) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint64_t z = (uint64_t)i + seed * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    dst[i] = src[z % src_words];
  }
}

__global__ void fill_kernel_u8(uint8_t* p, uint8_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_kernel_u32(uint32_t* p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_kernel_u64(uint64_t* p, uint64_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
} // namespace

namespace sg {

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int alloc_device(sg_device_t dev, void** p, size_t bytes) {
  (void)dev;
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return SG_ERR_NOMEM;
  }
  return SG_OK;
}

/* Zero with a kernel rather than hipMemsetAsync so the stream can be captured
 * into a HIP graph without memset nodes. */
int zero_async(sg_device_t dev, void* p, size_t bytes) {
  if (bytes == 0) return SG_OK;
  if ((bytes % 8) == 0 && (reinterpret_cast<uintptr_t>(p) % 8) == 0) {
    size_t n = bytes / 8;
    fill_kernel_u64<<<grid_for(dev, (int64_t)n, 256, 4), 256, 0, dev->stream>>>((uint64_t*)p, 0ull, n);
  } else {
    fill_kernel_u8<<<grid_for(dev, (int64_t)bytes, 256, 4), 256, 0, dev->stream>>>((uint8_t*)p, 0, bytes);
  }
  SG_HIP(hipGetLastError());
  return SG_OK;
}

int wait_host_flag(const int32_t* flag, int32_t want, hipStream_t s, const char* who) {
  for (uint64_t it = 1;; ++it) {
    if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == want) return SG_OK;
    if ((it & 4095) == 0) {
      const hipError_t q = hipStreamQuery(s);
      if (q == hipSuccess) {
        if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == want) return SG_OK;
        set_error("%s: the stream finished without the record %d", who, want);
        return SG_ERR_STATE;
      }
      if (q != hipErrorNotReady) SG_HIP(q);
    }
    __builtin_ia32_pause();
  }
}

} // namespace sg

namespace sg {
void device_release(sg_device_t dev) {
  if (dev->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  (void)hipSetDevice(dev->ordinal);
  if (dev->stream) {
    (void)hipStreamSynchronize(dev->stream);
    (void)hipStreamDestroy(dev->stream);
  }
  if (dev->host_scratch) (void)hipHostFree(dev->host_scratch);
  delete dev;
}
} // namespace sg

extern "C" {

const char* sg_last_error(void) { return g_last_error.c_str(); }
int sg_abi_version(void) { return SG_ABI_VERSION; }

int sg_device_count(int* count) {
  SG_CHECK(count, SG_ERR_INVALID, "sg_device_count: null count");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  return SG_OK;
}

int sg_device_open(int ordinal, sg_device_t* out) {
  SG_CHECK(out, SG_ERR_INVALID, "sg_device_open: null out");
  *out = nullptr;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c == 0) {
    sg::set_error("no HIP device available");
    return SG_ERR_NODEVICE;
  }
  SG_CHECK(ordinal >= 0 && ordinal < c, SG_ERR_INVALID, "sg_device_open: ordinal %d out of range [0,%d)", ordinal, c);
  SG_HIP(hipSetDevice(ordinal));
  auto* d = new sg_device_s();
  d->ordinal = ordinal;
  hipError_t e = hipGetDeviceProperties(&d->prop, ordinal);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc(&d->host_scratch, 1 << 20, hipHostMallocDefault);
  if (e != hipSuccess) {
    sg::set_error("sg_device_open: %s", hipGetErrorString(e));
    delete d;
    return SG_ERR_HIP;
  }
  d->host_scratch_bytes = 1 << 20;
  d->cus = d->prop.multiProcessorCount;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ordinal) != hipSuccess || khz <= 0) khz = 100000;
  d->wall_khz = khz;
  d->wait_ticks = (uint64_t)khz * 1000ull * 30ull; /* 30 s */
  sg::cc_scratch_init(d);
  *out = d;
  return SG_OK;
}

int sg_device_set_wait_limit(sg_device_t dev, double seconds) {
  SG_CHECK(dev, SG_ERR_INVALID, "sg_device_set_wait_limit: null device");
  SG_CHECK(seconds > 0.0 && seconds <= 3600.0, SG_ERR_INVALID, "sg_device_set_wait_limit: %g s outside (0, 3600]",
           seconds);
  dev->wait_ticks = (uint64_t)(seconds * 1000.0 * (double)dev->wall_khz);
  if (dev->wait_ticks == 0) dev->wait_ticks = 1;
  return SG_OK;
}

int sg_device_close(sg_device_t dev) {
  if (dev) sg::device_release(dev);
  return SG_OK;
}

int sg_device_info_get(sg_device_t dev, sg_device_info* info) {
  SG_CHECK(dev && info, SG_ERR_INVALID, "sg_device_info_get: null argument");
  std::memset(info, 0, sizeof(*info));
  std::strncpy(info->name, dev->prop.name, sizeof(info->name) - 1);
  std::strncpy(info->arch, dev->prop.gcnArchName, sizeof(info->arch) - 1);
  info->compute_units = dev->prop.multiProcessorCount;
  info->wave_size = dev->prop.warpSize;
  info->max_workgroup = dev->prop.maxThreadsPerBlock;
  info->max_waves_per_wg = dev->prop.maxThreadsPerBlock / (dev->prop.warpSize > 0 ? dev->prop.warpSize : 64);
  info->lds_bytes = (int64_t)dev->prop.sharedMemPerBlock;
  info->global_mem_bytes = (int64_t)dev->prop.totalGlobalMem;
  info->l2_bytes = (int64_t)dev->prop.l2CacheSize;
  info->clock_khz = dev->prop.clockRate;
  info->memory_clock_khz = dev->prop.memoryClockRate;
  info->memory_bus_bits = dev->prop.memoryBusWidth;
  info->device_ordinal = dev->ordinal;
  return SG_OK;
}

int sg_device_stream(sg_device_t dev, void** stream) {
  SG_CHECK(dev && stream, SG_ERR_INVALID, "sg_device_stream: null argument");
  *stream = (void*)dev->stream;
  return SG_OK;
}

int sg_device_synchronize(sg_device_t dev) {
  SG_CHECK(dev, SG_ERR_INVALID, "sg_device_synchronize: null device");
  SG_HIP(hipStreamSynchronize(dev->stream));
  return SG_OK;
}

int sg_alloc(sg_device_t dev, size_t bytes, int space, void** out) {
  SG_CHECK(dev && out, SG_ERR_INVALID, "sg_alloc: null argument");
  *out = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipSuccess;
  switch (space) {
    case SG_SPACE_DEVICE: e = hipMalloc(out, bytes); break;
    case SG_SPACE_SHARED: e = hipMallocManaged(out, bytes, hipMemAttachGlobal); break;
    case SG_SPACE_HOST: e = hipHostMalloc(out, bytes, hipHostMallocDefault); break;
    default: sg::set_error("sg_alloc: unknown memory space %d", space); return SG_ERR_INVALID;
  }
  if (e != hipSuccess) {
    sg::set_error("sg_alloc(%zu, space %d) failed: %s", bytes, space, hipGetErrorString(e));
    *out = nullptr;
    return SG_ERR_NOMEM;
  }
  return SG_OK;
}

int sg_free(sg_device_t dev, void* ptr) {
  (void)dev;
  if (!ptr) return SG_OK;
  hipPointerAttribute_t attr{};
  hipError_t e = hipPointerGetAttributes(&attr, ptr);
  if (e == hipSuccess && attr.type == hipMemoryTypeHost) {
    SG_HIP(hipHostFree(ptr));
  } else {
    SG_HIP(hipFree(ptr));
  }
  return SG_OK;
}

int sg_memcpy(sg_device_t dev, void* dst, const void* src, size_t bytes) {
  SG_CHECK(dev, SG_ERR_INVALID, "sg_memcpy: null device");
  if (bytes == 0) return SG_OK;
  SG_CHECK(dst && src, SG_ERR_INVALID, "sg_memcpy: null pointer");
  SG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, dev->stream));
  SG_HIP(hipStreamSynchronize(dev->stream));
  return SG_OK;
}

int sg_prefetch(sg_device_t dev, const void* ptr, size_t bytes) {
  SG_CHECK(dev, SG_ERR_INVALID, "sg_prefetch: null device");
  if (bytes == 0) return SG_OK;
  SG_CHECK(ptr, SG_ERR_INVALID, "sg_prefetch: null pointer");
  const hipError_t e = hipMemPrefetchAsync(ptr, bytes, dev->ordinal, dev->stream);
  if (e != hipSuccess) {
    (void)hipGetLastError(); /* not migratable (not managed, or no support): a hint only */
    return SG_OK;
  }
  SG_HIP(hipStreamSynchronize(dev->stream));
  return SG_OK;
}

int sg_copy_device(sg_device_t dev, void* dst, const void* src, size_t bytes) {
  SG_CHECK(dev, SG_ERR_INVALID, "sg_copy_device: null device");
  if (bytes == 0) return SG_OK;
  SG_CHECK(dst && src, SG_ERR_INVALID, "sg_copy_device: null pointer");
  SG_CHECK(((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0, SG_ERR_INVALID,
           "sg_copy_device: pointers and size must be multiples of 16 bytes");
  const size_t n16 = bytes / 16;
  copy_kernel_u128<<<sg::grid_for(dev, (int64_t)n16, 256, 8), 256, 0, dev->stream>>>((const uint4*)src, (uint4*)dst, n16);
  SG_HIP(hipGetLastError());
  SG_HIP(hipStreamSynchronize(dev->stream));
  return SG_OK;
}

int sg_gather_device(sg_device_t dev, uint32_t* dst, const uint32_t* src, uint64_t src_words, size_t count,
                     uint64_t seed) {
  SG_CHECK(dev, SG_ERR_INVALID, "sg_gather_device: null device");
  if (count == 0) return SG_OK;
  SG_CHECK(dst && src && src_words > 0, SG_ERR_INVALID, "sg_gather_device: null pointer or empty source");
  gather_kernel_u32<<<sg::grid_for(dev, (int64_t)count, 256, 8), 256, 0, dev->stream>>>(src, src_words, dst, count,
                                                                                      seed);
  SG_HIP(hipGetLastError());
  SG_HIP(hipStreamSynchronize(dev->stream));
  return SG_OK;
}

int sg_fill(sg_device_t dev, void* dst, uint64_t pattern, int pattern_bytes, size_t count) {
  SG_CHECK(dev, SG_ERR_INVALID, "sg_fill: null device");
  if (count == 0) return SG_OK;
  SG_CHECK(dst, SG_ERR_INVALID, "sg_fill: null pointer");
  int g = sg::grid_for(dev, (int64_t)count, 256, 4);
  switch (pattern_bytes) {
    case 1: fill_kernel_u8<<<g, 256, 0, dev->stream>>>((uint8_t*)dst, (uint8_t)pattern, count); break;
    case 4: fill_kernel_u32<<<g, 256, 0, dev->stream>>>((uint32_t*)dst, (uint32_t)pattern, count); break;
    case 8: fill_kernel_u64<<<g, 256, 0, dev->stream>>>((uint64_t*)dst, pattern, count); break;
    default: sg::set_error("sg_fill: pattern_bytes must be 1, 4 or 8 (got %d)", pattern_bytes); return SG_ERR_INVALID;
  }
  SG_HIP(hipGetLastError());
  SG_HIP(hipStreamSynchronize(dev->stream));
  return SG_OK;
}

int sg_event_create(sg_device_t dev, sg_event_t* out) {
  SG_CHECK(dev && out, SG_ERR_INVALID, "sg_event_create: null argument");
  auto* e = new sg_event_s();
  e->dev = dev;
  hipError_t r = hipEventCreate(&e->ev);
  if (r != hipSuccess) {
    delete e;
    sg::set_error("hipEventCreate: %s", hipGetErrorString(r));
    return SG_ERR_HIP;
  }
  *out = e;
  return SG_OK;
}

int sg_event_destroy(sg_event_t ev) {
  if (!ev) return SG_OK;
  (void)hipEventDestroy(ev->ev);
  delete ev;
  return SG_OK;
}

int sg_event_record(sg_event_t ev) {
  SG_CHECK(ev, SG_ERR_INVALID, "sg_event_record: null event");
  SG_HIP(hipEventRecord(ev->ev, ev->dev->stream));
  return SG_OK;
}

int sg_event_synchronize(sg_event_t ev) {
  SG_CHECK(ev, SG_ERR_INVALID, "sg_event_synchronize: null event");
  SG_HIP(hipEventSynchronize(ev->ev));
  return SG_OK;
}

int sg_event_elapsed_ms(sg_event_t start, sg_event_t stop, float* ms) {
  SG_CHECK(start && stop && ms, SG_ERR_INVALID, "sg_event_elapsed_ms: null argument");
  SG_HIP(hipEventSynchronize(stop->ev));
  SG_HIP(hipEventElapsedTime(ms, start->ev, stop->ev));
  return SG_OK;
}

} // extern "C"
