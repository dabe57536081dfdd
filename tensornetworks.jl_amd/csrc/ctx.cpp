// Context, device buffers and host<->device transfers of libtnhip.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <sstream>

#include "tnh_internal.hpp"

// ------------------------------------------------------------------ allocator

void *tnh_ctx::alloc(size_t bytes, size_t *cap_out) {
  if (arena) throw tnh::Error(TNH_ERR_NOMEM, "device allocation requested inside an entry point that allocates nothing");
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~size_t(255);
  // reuse a cached block of size in [bytes, 2*bytes]
  auto it = cache.lower_bound(bytes);
  if (it != cache.end() && it->first <= 2 * bytes) {
    void *p = it->second;
    *cap_out = it->first;
    cached_bytes -= it->first;
    cache.erase(it);
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw tnh::Error(TNH_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  *cap_out = bytes;
  return p;
}

void tnh_ctx::release(void *p, size_t cap) {
  if (!p) return;
  cache.emplace(cap, p);
  cached_bytes += cap;
  // keep the cache bounded: beyond 8 GiB, free the largest blocks
  const size_t kMaxCache = size_t(8) << 30;
  while (cached_bytes > kMaxCache && !cache.empty()) {
    (void)hipStreamSynchronize(stream);
    auto last = std::prev(cache.end());
    (void)hipFree(last->second);
    cached_bytes -= last->first;
    cache.erase(last);
  }
}

void tnh_ctx::trim() {
  (void)hipStreamSynchronize(stream);
  for (auto &kv : cache) (void)hipFree(kv.second);
  cache.clear();
  cached_bytes = 0;
}

void *tnh_ctx::workspace(int slot, size_t bytes) {
  if (arena) {
    auto it = arena->slots.find(slot);
    if (it != arena->slots.end() && it->second.second >= bytes) return it->second.first;
    const size_t need = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
    if (arena->off + need > arena->cap)
      throw tnh::Error(TNH_ERR_NOMEM, "static scratch exhausted (" + std::to_string(arena->off + need) +
                                          " bytes needed, " + std::to_string(arena->cap) + " available)");
    void *p = arena->base + arena->off;
    arena->off += need;
    arena->slots[slot] = {p, need};
    return p;
  }
  auto &w = ws[slot];
  if (w.second < bytes) {
    if (w.first) release(w.first, w.second);
    size_t cap = 0;
    w.first = alloc(bytes, &cap);
    w.second = cap;
  }
  return w.first;
}

hipEvent_t tnh_ctx::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  TNH_HIP(hipEventCreate(&e));
  return e;
}

hipEvent_t tnh_ctx::prof_begin() {
  if (!profiling) return nullptr;
  hipEvent_t e = take_event();
  TNH_HIP(hipEventRecord(e, stream));
  return e;
}

void tnh_ctx::prof_end(const char *family, hipEvent_t start, double flops, double bytes) {
  if (!profiling || !start) return;
  hipEvent_t e = take_event();
  TNH_HIP(hipEventRecord(e, stream));
  prof_pending[family].push_back({start, e, flops, bytes});
  // keep the number of outstanding events bounded
  size_t n = 0;
  for (auto &kv : prof_pending) n += kv.second.size();
  if (n > 4096) prof_resolve();
}

void tnh_ctx::prof_resolve() {
  if (prof_pending.empty()) return;
  TNH_HIP(hipStreamSynchronize(stream));
  for (auto &kv : prof_pending) {
    ProfSum &s = prof_done[kv.first];
    for (auto &r : kv.second) {
      float ms = 0;
      TNH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      s.launches += 1;
      s.ms += ms;
      s.flops += r.flops;
      s.bytes += r.bytes;
      event_pool.push_back(r.a);
      event_pool.push_back(r.b);
    }
  }
  prof_pending.clear();
}

hipStream_t tnh_ctx::aux_stream() {
  if (!aux) TNH_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
  return aux;
}

tnh_ctx::~tnh_ctx() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (aux) (void)hipStreamSynchronize(aux);
  for (auto &kv : prof_pending)
    for (auto &r : kv.second) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
  for (auto e : event_pool) (void)hipEventDestroy(e);
  for (auto &kv : ws)
    if (kv.second.first) (void)hipFree(kv.second.first);
  for (auto &kv : cache) (void)hipFree(kv.second);
  for (auto &kv : schedules)
    if (kv.second.first) (void)hipFree(kv.second.first);
  if (host_scalars) (void)hipHostFree(host_scalars);
  if (pin_down) (void)hipHostFree(pin_down);
  if (pin_up) (void)hipHostFree(pin_up);
  if (up_done) (void)hipEventDestroy(up_done);
  if (aux) (void)hipStreamDestroy(aux);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace tnh {

tnh_buf *buf_new(tnh_ctx *ctx, int dtype, int ndim, const int64_t *dims) {
  TNH_REQUIRE(valid_dtype(dtype), TNH_ERR_ARG, "invalid dtype " + std::to_string(dtype));
  TNH_REQUIRE(ndim >= 0 && ndim <= 8, TNH_ERR_ARG, "ndim must be in [0, 8]");
  auto *b = new tnh_buf();
  b->dtype = dtype;
  b->ndim = ndim;
  int64_t n = 1;
  for (int i = 0; i < ndim; ++i) {
    TNH_REQUIRE(dims[i] >= 0, TNH_ERR_ARG, "negative dimension");
    b->dims[i] = dims[i];
    n *= dims[i];
  }
  b->numel = n;
  b->bytes = size_t(n) * real_size(dtype) * (is_complex(dtype) ? 2 : 1);
  try {
    b->ptr = ctx->alloc(b->bytes, &b->cap);
  } catch (...) {
    delete b;
    throw;
  }
  return b;
}

void buf_delete(tnh_ctx *ctx, tnh_buf *b) {
  if (!b) return;
  if (b->owned) ctx->release(b->ptr, b->cap);
  delete b;
}

void buf_ensure(tnh_ctx *ctx, tnh_buf **b, int dtype, int ndim, const int64_t *dims) {
  if (*b) {
    bool same = (*b)->dtype == dtype && (*b)->ndim == ndim;
    for (int i = 0; same && i < ndim; ++i) same = (*b)->dims[i] == dims[i];
    if (same) return;
    buf_delete(ctx, *b);
    *b = nullptr;
  }
  *b = buf_new(ctx, dtype, ndim, dims);
}

std::string shape_str(const tnh_buf *b) {
  std::ostringstream s;
  s << "(";
  for (int i = 0; i < b->ndim; ++i) s << (i ? "," : "") << b->dims[i];
  s << ")";
  return s.str();
}

void to_host(tnh_ctx *ctx, void *host, const void *dev, size_t bytes) {
  if (bytes == 0) return;
  if (ctx->pin_down_cap < bytes) {
    if (ctx->pin_down) TNH_HIP(hipHostFree(ctx->pin_down));
    ctx->pin_down = nullptr;
    ctx->pin_down_cap = 0;
    const size_t cap = std::max<size_t>(bytes, size_t(1) << 16);
    TNH_HIP(hipHostMalloc(&ctx->pin_down, cap, hipHostMallocDefault));
    ctx->pin_down_cap = cap;
  }
  TNH_HIP(hipMemcpyAsync(ctx->pin_down, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
  TNH_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(host, ctx->pin_down, bytes);
  // every queued upload has completed: the upload arena is free again
  ctx->pin_up_off = 0;
  ctx->up_pending = false;
}

void to_device(tnh_ctx *ctx, void *dev, const void *host, size_t bytes) {
  if (bytes == 0) return;
  const size_t need = (bytes + 255) & ~size_t(255);
  if (ctx->pin_up_off + need > ctx->pin_up_cap) {
    // wait until the queued uploads have left the arena, then start over (growing it)
    if (ctx->up_pending) TNH_HIP(hipEventSynchronize(ctx->up_done));
    ctx->up_pending = false;
    ctx->pin_up_off = 0;
    if (ctx->pin_up_cap < need) {
      if (ctx->pin_up) TNH_HIP(hipHostFree(ctx->pin_up));
      ctx->pin_up = nullptr;
      ctx->pin_up_cap = 0;
      const size_t cap = std::max<size_t>(need, size_t(1) << 20);
      TNH_HIP(hipHostMalloc(&ctx->pin_up, cap, hipHostMallocDefault));
      ctx->pin_up_cap = cap;
    }
  }
  char *src = static_cast<char *>(ctx->pin_up) + ctx->pin_up_off;
  std::memcpy(src, host, bytes);
  TNH_HIP(hipMemcpyAsync(dev, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  if (!ctx->up_done) TNH_HIP(hipEventCreateWithFlags(&ctx->up_done, hipEventDisableTiming));
  TNH_HIP(hipEventRecord(ctx->up_done, ctx->stream));
  ctx->up_pending = true;
  ctx->pin_up_off += need;
}

}  // namespace tnh

using namespace tnh;

extern "C" {

const char *tnh_version(void) { return "tnhip 0.1 gfx950"; }

int tnh_ctx_create(int device, tnh_ctx **out) {
  if (!out) return TNH_ERR_ARG;
  *out = nullptr;
  auto *ctx = new (std::nothrow) tnh_ctx();
  if (!ctx) return TNH_ERR_NOMEM;
  int rc = tnh_guard(ctx, [&] {
    int n = 0;
    TNH_HIP(hipGetDeviceCount(&n));
    TNH_REQUIRE(device >= 0 && device < n, TNH_ERR_ARG,
                "device " + std::to_string(device) + " out of range (" + std::to_string(n) +
                    " devices)");
    TNH_HIP(hipSetDevice(device));
    ctx->device = device;
    hipDeviceProp_t prop;
    TNH_HIP(hipGetDeviceProperties(&prop, device));
    ctx->cu_count = prop.multiProcessorCount;
    TNH_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    // this device's copy of the static scratch of tnh_heff_apply_noalloc (once per device)
    noalloc_init(device);
    TNH_HIP(hipHostMalloc((void **)&ctx->host_scalars, sizeof(double) * tnh_ctx::kHostScalars,
                          hipHostMallocDefault));
  });
  if (rc != TNH_OK) {
    // keep the message readable by the caller through a static copy
    static thread_local std::string last;
    last = ctx->err;
    std::fprintf(stderr, "tnh_ctx_create: %s\n", last.c_str());
    delete ctx;
    return rc;
  }
  *out = ctx;
  return TNH_OK;
}

int tnh_ctx_destroy(tnh_ctx *ctx) {
  if (!ctx) return TNH_OK;
  delete ctx;
  return TNH_OK;
}

const char *tnh_last_error(tnh_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tnh_ctx_synchronize(tnh_ctx *ctx) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] { TNH_HIP(hipStreamSynchronize(ctx->stream)); });
}

int tnh_profile_enable(tnh_ctx *ctx, int on) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    if (!on) ctx->prof_resolve();
    ctx->profiling = on != 0;
  });
}

int tnh_profile_reset(tnh_ctx *ctx) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    ctx->prof_resolve();
    ctx->prof_done.clear();
  });
}

int tnh_profile_read(tnh_ctx *ctx, const char *family, int64_t *launches, double *total_ms,
                     double *flops, double *bytes) {
  if (!ctx || !family) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    ctx->prof_resolve();
    auto it = ctx->prof_done.find(family);
    tnh_ctx::ProfSum s;
    if (it != ctx->prof_done.end()) s = it->second;
    if (launches) *launches = s.launches;
    if (total_ms) *total_ms = s.ms;
    if (flops) *flops = s.flops;
    if (bytes) *bytes = s.bytes;
  });
}

int tnh_buf_create(tnh_ctx *ctx, int dtype, int ndim, const int64_t *dims, tnh_buf **out) {
  if (!ctx || !out) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(ndim == 0 || dims, TNH_ERR_ARG, "dims is NULL");
    *out = buf_new(ctx, dtype, ndim, dims);
  });
}

int tnh_buf_upload(tnh_ctx *ctx, const void *host, int host_dtype, int dtype, int ndim,
                   const int64_t *dims, tnh_buf **out) {
  if (!ctx || !out) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(valid_dtype(host_dtype), TNH_ERR_ARG, "invalid host dtype");
    TNH_REQUIRE(ndim == 0 || dims, TNH_ERR_ARG, "dims is NULL");
    tnh_buf *b = buf_new(ctx, dtype, ndim, dims);
    try {
      int64_t n = b->numel;
      TNH_REQUIRE(n == 0 || host, TNH_ERR_ARG, "host pointer is NULL");
      if (n > 0) {
        if (is_complex(host_dtype) && !is_complex(dtype)) {
          // a real device tensor may only receive a complex array with zero imaginary parts
          bool nonzero = false;
          if (host_dtype == TNH_C128) {
            const double *h = static_cast<const double *>(host);
            for (int64_t i = 0; i < n && !nonzero; ++i) nonzero = h[2 * i + 1] != 0.0;
          } else {
            const float *h = static_cast<const float *>(host);
            for (int64_t i = 0; i < n && !nonzero; ++i) nonzero = h[2 * i + 1] != 0.0f;
          }
          TNH_REQUIRE(!nonzero, TNH_ERR_NUMERIC,
                      "complex array with non-zero imaginary part uploaded to a real tensor");
        }
        size_t hbytes = size_t(n) * real_size(host_dtype) * (is_complex(host_dtype) ? 2 : 1);
        bool direct = host_dtype == dtype && !is_complex(dtype);
        if (direct) {
          TNH_HIP(hipMemcpyAsync(b->ptr, host, hbytes, hipMemcpyHostToDevice, ctx->stream));
        } else {
          void *stage = ctx->workspace(/*slot*/ 900, hbytes);
          TNH_HIP(hipMemcpyAsync(stage, host, hbytes, hipMemcpyHostToDevice, ctx->stream));
          reformat(ctx, stage, host_dtype, true, b->ptr, dtype, false, n);
        }
        // the host buffer is borrowed only for this call
        TNH_HIP(hipStreamSynchronize(ctx->stream));
      }
    } catch (...) {
      buf_delete(ctx, b);
      throw;
    }
    *out = b;
  });
}

int tnh_buf_download(tnh_ctx *ctx, const tnh_buf *b, void *host, int host_dtype) {
  if (!ctx || !b) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(valid_dtype(host_dtype), TNH_ERR_ARG, "invalid host dtype");
    int64_t n = b->numel;
    if (n == 0) return;
    TNH_REQUIRE(host, TNH_ERR_ARG, "host pointer is NULL");
    size_t hbytes = size_t(n) * real_size(host_dtype) * (is_complex(host_dtype) ? 2 : 1);
    if (host_dtype == b->dtype && !is_complex(b->dtype)) {
      to_host(ctx, host, b->ptr, hbytes);
    } else {
      void *stage = ctx->workspace(901, hbytes);
      reformat(ctx, b->ptr, b->dtype, false, stage, host_dtype, true, n);
      to_host(ctx, host, stage, hbytes);
    }
  });
}

int tnh_buf_wrap(tnh_ctx *ctx, void *device_ptr, int dtype, int ndim, const int64_t *dims,
                 tnh_buf **out) {
  if (!ctx || !out) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(valid_dtype(dtype), TNH_ERR_ARG, "invalid dtype");
    TNH_REQUIRE(ndim >= 0 && ndim <= 8, TNH_ERR_ARG, "ndim must be in [0, 8]");
    auto *b = new tnh_buf();
    b->dtype = dtype;
    b->ndim = ndim;
    int64_t n = 1;
    for (int i = 0; i < ndim; ++i) {
      b->dims[i] = dims[i];
      n *= dims[i];
    }
    b->numel = n;
    b->bytes = size_t(n) * real_size(dtype) * (is_complex(dtype) ? 2 : 1);
    b->cap = b->bytes;
    b->ptr = device_ptr;
    b->owned = false;
    *out = b;
  });
}

int tnh_buf_free(tnh_ctx *ctx, tnh_buf *b) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] { buf_delete(ctx, b); });
}

int tnh_buf_shape(const tnh_buf *b, int *dtype, int *ndim, int64_t *dims) {
  if (!b) return TNH_ERR_ARG;
  if (dtype) *dtype = b->dtype;
  if (ndim) *ndim = b->ndim;
  if (dims)
    for (int i = 0; i < 8; ++i) dims[i] = i < b->ndim ? b->dims[i] : 1;
  return TNH_OK;
}

void *tnh_buf_device_ptr(const tnh_buf *b) { return b ? b->ptr : nullptr; }

int tnh_buf_convert(tnh_ctx *ctx, const tnh_buf *src, int dtype, tnh_buf **out) {
  if (!ctx || !src || !out) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    tnh_buf *b = buf_new(ctx, dtype, src->ndim, src->dims);
    if (src->numel > 0) reformat(ctx, src->ptr, src->dtype, false, b->ptr, dtype, false, src->numel);
    *out = b;
  });
}

}  // extern "C"
