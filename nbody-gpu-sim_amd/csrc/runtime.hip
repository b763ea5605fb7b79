// Runtime half of the C ABI: status strings, last-error slot, memory and
// stream helpers, and the ParticleDataManager equivalents
// (src/cuda/particle_init.cu:11-160).
#include <cstring>
#include <mutex>

#include "common.hpp"

namespace nbx {

static thread_local char g_last_error[512] = "";

int fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return status;
}

void clear_error() { g_last_error[0] = '\0'; }

int Scratch::ensure(size_t need) {
  if (need <= bytes) return NBX_OK;
  if (ptr) {
    hipError_t e = hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (e != hipSuccess) return fail(NBX_ERR_HIP, "hipFree: %s", hipGetErrorString(e));
  }
  size_t grow = need + need / 4;
  if (hipMalloc(&ptr, grow) != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
    return fail(NBX_ERR_NOMEM, "device scratch allocation of %zu bytes failed", grow);
  }
  bytes = grow;
  return NBX_OK;
}

Scratch::~Scratch() {
  // Process teardown may already have unloaded the runtime; a failure here is ignored.
  if (ptr) (void)hipFree(ptr);
}

}  // namespace nbx

using namespace nbx;

extern "C" {

const char* nbx_status_string(int status) {
  switch (status) {
    case NBX_OK: return "ok";
    case NBX_ERR_INVALID: return "invalid argument";
    case NBX_ERR_HIP: return "HIP runtime error";
    case NBX_ERR_NOMEM: return "out of memory";
    case NBX_ERR_CAPACITY: return "capacity exceeded";
    case NBX_ERR_STATE: return "invalid call order";
    default: return "unknown status";
  }
}

const char* nbx_last_error(void) { return g_last_error; }

int nbx_device_info(char* name, size_t name_len, int* cu_count, int* max_clock_khz,
                    size_t* total_mem) {
  hipDeviceProp_t prop;
  NBX_HIP(hipGetDeviceProperties(&prop, 0));
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (max_clock_khz) *max_clock_khz = prop.clockRate;
  if (total_mem) *total_mem = prop.totalGlobalMem;
  return NBX_OK;
}

int nbx_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(NBX_ERR_INVALID, "nbx_malloc: null output");
  *ptr = nullptr;
  if (bytes == 0) return NBX_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(NBX_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return NBX_OK;
}

int nbx_free(void* ptr) {
  if (ptr) NBX_HIP(hipFree(ptr));
  return NBX_OK;
}

int nbx_memcpy_htod(void* dst, const void* src, size_t bytes) {
  if (bytes) NBX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return NBX_OK;
}
int nbx_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  if (bytes) NBX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return NBX_OK;
}
int nbx_memcpy_dtod(void* dst, const void* src, size_t bytes) {
  if (bytes) NBX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return NBX_OK;
}
int nbx_synchronize(void) {
  NBX_HIP(hipDeviceSynchronize());
  return NBX_OK;
}
int nbx_stream_create(nbx_stream* out) {
  if (!out) return fail(NBX_ERR_INVALID, "nbx_stream_create: null output");
  hipStream_t s;
  NBX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = s;
  return NBX_OK;
}
int nbx_stream_destroy(nbx_stream s) {
  if (s) NBX_HIP(hipStreamDestroy(as_stream(s)));
  return NBX_OK;
}
int nbx_stream_synchronize(nbx_stream s) {
  NBX_HIP(hipStreamSynchronize(as_stream(s)));
  return NBX_OK;
}

int nbx_event_create(void** ev) {
  if (!ev) return fail(NBX_ERR_INVALID, "nbx_event_create: null output");
  hipEvent_t e;
  NBX_HIP(hipEventCreate(&e));
  *ev = e;
  return NBX_OK;
}
int nbx_event_destroy(void* ev) {
  if (ev) NBX_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)));
  return NBX_OK;
}
int nbx_event_record(void* ev, nbx_stream s) {
  NBX_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), as_stream(s)));
  return NBX_OK;
}
int nbx_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (!ms) return fail(NBX_ERR_INVALID, "nbx_event_elapsed_ms: null output");
  NBX_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
  NBX_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                              reinterpret_cast<hipEvent_t>(stop)));
  return NBX_OK;
}

static float** soa_field(nbx_particles* p, int k) {
  float** f[13] = {&p->pos_x, &p->pos_y, &p->pos_z, &p->vel_x, &p->vel_y,
                   &p->vel_z, &p->acc_x, &p->acc_y, &p->acc_z, &p->acc_old_x,
                   &p->acc_old_y, &p->acc_old_z, &p->mass};
  return f[k];
}

int nbx_particles_free(nbx_particles* p) {
  if (!p) return NBX_OK;
  int rc = NBX_OK;
  for (int k = 0; k < 13; ++k) {
    float** f = soa_field(p, k);
    if (*f && hipFree(*f) != hipSuccess && rc == NBX_OK)
      rc = fail(NBX_ERR_HIP, "hipFree failed while releasing particle arrays");
    *f = nullptr;
  }
  p->count = 0;
  return rc;
}

int nbx_particles_alloc(nbx_particles* p, size_t count) {
  if (!p) return fail(NBX_ERR_INVALID, "nbx_particles_alloc: null struct");
  memset(p, 0, sizeof(*p));
  p->count = count;
  const size_t bytes = count * sizeof(float);
  // One allocation per array keeps the reference's 13-pointer layout; each
  // array starts 256-byte aligned (hipMalloc), which the vector loads rely on.
  for (int k = 0; k < 13; ++k) {
    float** f = soa_field(p, k);
    if (bytes && hipMalloc(reinterpret_cast<void**>(f), bytes) != hipSuccess) {
      (void)hipGetLastError();
      *f = nullptr;
      nbx_particles_free(p);
      return fail(NBX_ERR_NOMEM, "device allocation of %zu particles failed", count);
    }
  }
  for (int k = 6; k < 12 && bytes; ++k) {
    if (hipMemset(*soa_field(p, k), 0, bytes) != hipSuccess) {
      nbx_particles_free(p);
      return fail(NBX_ERR_HIP, "hipMemset of acceleration arrays failed");
    }
  }
  return NBX_OK;
}

static const int kCopied[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 12};

int nbx_particles_upload(nbx_particles* dev, const nbx_particles* host) {
  if (!dev || !host) return fail(NBX_ERR_INVALID, "nbx_particles_upload: null struct");
  if (host->count > dev->count)
    return fail(NBX_ERR_CAPACITY, "upload of %zu particles into %zu", host->count, dev->count);
  const size_t bytes = host->count * sizeof(float);
  nbx_particles h = *host;
  for (int k : kCopied)
    if (bytes) NBX_HIP(hipMemcpy(*soa_field(dev, k), *soa_field(&h, k), bytes, hipMemcpyHostToDevice));
  return NBX_OK;
}

int nbx_particles_download(nbx_particles* host, const nbx_particles* dev) {
  if (!dev || !host) return fail(NBX_ERR_INVALID, "nbx_particles_download: null struct");
  if (dev->count > host->count)
    return fail(NBX_ERR_CAPACITY, "download of %zu particles into %zu", dev->count, host->count);
  const size_t bytes = dev->count * sizeof(float);
  nbx_particles d = *dev;
  for (int k : kCopied)
    if (bytes) NBX_HIP(hipMemcpy(*soa_field(host, k), *soa_field(&d, k), bytes, hipMemcpyDeviceToHost));
  return NBX_OK;
}

}  // extern "C"
