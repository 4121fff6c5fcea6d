// mpk_runtime.hip -- context, memory and error plumbing behind the C ABI.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "mpk_internal.h"
#include "../../include/mpk.h"

namespace mpk {

static Ctx g_ctx;
static bool g_init_done = false;

// The scratch of the no-allocation entry point (mpk_vec_dot_noalloc): a
// fixed array in the library's code object, 65,536 doubles (512 KB) on each
// GPU the library runs on, held until the process ends.  Host code never uses
// the name `scratch` as an address (on the host it is only a stand-in): init
// looks up the device address once per device, with that device current, and
// everything else uses that address.
#define MPK_FIXED_SCRATCH 65536
static __device__ __align__(256) double scratch[MPK_FIXED_SCRATCH];
#define MPK_MAX_DEVICES 64
static double *g_fixed_ptr[MPK_MAX_DEVICES] = {};
static std::mutex g_mu;
static int g_status = 0;
static char g_msg[1024] = "";
// Polls a sweep's wait makes before it gives up: at least ~1 s at one poll
// per sleep-and-load round trip, far beyond any wait of a well-formed plan
// (a whole sweep takes milliseconds).
#define MPK_WATCH_DEFAULT_POLLS (1u << 24)
static std::atomic<unsigned> g_watch_limit{MPK_WATCH_DEFAULT_POLLS};
static std::atomic<unsigned> g_watch_gen{1};

void set_error(int code, const char *fmt, ...) {
  if (g_status != 0) return;  // keep the first error
  g_status = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
}

bool failed() { return g_status != 0; }

static bool init_locked(int device) {
  if (g_init_done) return g_ctx.device >= 0;
  g_init_done = true;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    set_error(2, "no HIP device is available (%s)", e != hipSuccess ? hipGetErrorString(e) : "count 0");
    return false;
  }
  if (device < 0) HIPCHK(hipGetDevice(&device));
  HIPCHK(hipSetDevice(device));
  // default priority: the stream is created without a priority argument
  HIPCHK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&g_ctx.dres, 1024 * sizeof(double)));
  HIPCHK(hipHostMalloc(&g_ctx.hres, 1024 * sizeof(double), hipHostMallocDefault));
  HIPCHK(hipMalloc(&g_ctx.tickets, 64 * sizeof(unsigned)));
  HIPCHK(hipMemset(g_ctx.tickets, 0, 64 * sizeof(unsigned)));
  // fine-grained (coherent) host memory: the host's abort store reaches a
  // running kernel, and a kernel's fault store reaches the host
  HIPCHK(hipHostMalloc(&g_ctx.watch, 4 * sizeof(unsigned), hipHostMallocCoherent));
  if (g_ctx.watch) std::memset(g_ctx.watch, 0, 4 * sizeof(unsigned));
  if (device >= 0 && device < MPK_MAX_DEVICES && !g_fixed_ptr[device]) {
    void *p = nullptr;
    hipError_t se = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !p) {
      set_error(1, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(se));
      return false;
    }
    g_fixed_ptr[device] = static_cast<double *>(p);
  }
  g_ctx.fixed = device >= 0 && device < MPK_MAX_DEVICES ? g_fixed_ptr[device] : nullptr;
  g_ctx.fixed_cap = g_ctx.fixed ? MPK_FIXED_SCRATCH : 0;
  HIPCHK(hipDeviceSynchronize());
  if (failed()) return false;
  g_ctx.device = device;
  return true;
}

Ctx *ctx_ready() { return g_init_done && g_ctx.device >= 0 ? &g_ctx : nullptr; }

Ctx *ctx() {
  if (g_init_done) return g_ctx.device >= 0 ? &g_ctx : nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return init_locked(-1) ? &g_ctx : nullptr;
}

hipStream_t stream() {
  Ctx *c = ctx();
  return c ? c->stream : nullptr;
}

double *partials(size_t need) {
  Ctx *c = ctx();
  if (!c) return nullptr;
  if (need > c->partials_cap) {
    if (c->partials) {
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipFree(c->partials));
    }
    size_t cap = need < 65536 ? 65536 : need;
    HIPCHK(hipMalloc(&c->partials, cap * sizeof(double)));
    c->partials_cap = cap;
  }
  return c->partials;
}

void copy_dtoh(void *dst, const void *src, size_t bytes) {
  if (!bytes || !ctx()) return;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
  HIPCHK(hipStreamSynchronize(stream()));
}
void copy_htod(void *dst, const void *src, size_t bytes) {
  if (!bytes || !ctx()) return;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
  HIPCHK(hipStreamSynchronize(stream()));
}

int *flag_buffer(size_t need) {
  Ctx *c = ctx();
  if (!c) return nullptr;
  if (need > c->flags_cap) {
    if (c->flags) {
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipFree(c->flags));
    }
    size_t cap = need < 65536 ? 65536 : need;
    HIPCHK(hipMalloc(&c->flags, cap * sizeof(int)));
    HIPCHK(hipMemset(c->flags, 0, cap * sizeof(int)));
    c->flags_cap = cap;
  }
  return c->flags;
}

double *scratch_vector(size_t need) {
  Ctx *c = ctx();
  if (!c) return nullptr;
  if (need > c->dtmp_cap) {
    if (c->dtmp) {
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipFree(c->dtmp));
    }
    size_t cap = need < 4096 ? 4096 : need;
    HIPCHK(hipMalloc(&c->dtmp, cap * sizeof(double)));
    c->dtmp_cap = cap;
  }
  return c->dtmp;
}

// Records an error when a kernel has raised the sweep watch's fault word
// since the last check (and clears the word).
static void watch_check() {
  Ctx *c = ctx_ready();
  if (!c || !c->watch) return;
  volatile unsigned *w = c->watch;
  if (w[1]) {
    w[1] = 0u;
    set_error(10, "a sweep's wait for a published row gave up (after %u polls, or on the abort word): "
                  "the sweep's results are not valid", g_watch_limit.load());
  }
}

void sync() {
  Ctx *c = ctx();
  if (!c) return;
  HIPCHK(hipStreamSynchronize(c->stream));
  watch_check();
}

void sweep_watch_state(unsigned **word, unsigned *limit, unsigned *gen) {
  Ctx *c = ctx_ready();
  *word = c ? c->watch : nullptr;
  *limit = g_watch_limit.load();
  *gen = c && c->watch ? g_watch_gen.load() : 0u;
}

// -1: not read yet; 0: tile-tree order; 1: the reference's order
static std::atomic<int> g_red_order{-1};

static int red_order_from_env() {
  const char *e = std::getenv("MPK_REDUCTION");
  return e && std::strcmp(e, "ref") == 0 ? 1 : 0;
}

bool red_ref() {
  int o = g_red_order.load(std::memory_order_relaxed);
  if (o < 0) {
    o = red_order_from_env();
    g_red_order.store(o, std::memory_order_relaxed);
  }
  return o == 1;
}

}  // namespace mpk

using namespace mpk;

extern "C" {

int64_t mpk_init(int64_t device) {
  std::lock_guard<std::mutex> lk(g_mu);
  return init_locked((int)device) ? 0 : (g_status ? g_status : 2);
}

int64_t mpk_status(void) {
  watch_check();
  return g_status;
}

int64_t mpk_last_error(int64_t buf, int64_t len) {
  if (buf && len > 0) {
    strncpy(reinterpret_cast<char *>(buf), g_msg, (size_t)len - 1);
    reinterpret_cast<char *>(buf)[len - 1] = 0;
  }
  return g_status;
}

void mpk_clear_error(void) {
  g_status = 0;
  g_msg[0] = 0;
}

int64_t mpk_device_name(int64_t buf, int64_t len) {
  Ctx *c = ctx();
  if (!c || !buf || len <= 0) return -1;
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, c->device));
  snprintf(reinterpret_cast<char *>(buf), (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName,
           p.multiProcessorCount);
  return 0;
}

int64_t mpk_malloc(int64_t bytes) {
  if (!ctx()) return 0;
  void *p = nullptr;
  size_t b = bytes > 0 ? (size_t)bytes : 16;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    set_error(3, "hipMalloc(%lld) failed: %s", (long long)bytes, hipGetErrorString(e));
    return 0;
  }
  return reinterpret_cast<int64_t>(p);
}

void mpk_free(int64_t ptr) {
  if (!ptr || !ctx()) return;
  HIPCHK(hipStreamSynchronize(stream()));
  HIPCHK(hipFree(reinterpret_cast<void *>(ptr)));
}

void mpk_memcpy_htod(int64_t dst, int64_t src, int64_t bytes) {
  if (bytes <= 0 || !ctx()) return;
  HIPCHK(hipMemcpyAsync(reinterpret_cast<void *>(dst), reinterpret_cast<const void *>(src),
                        (size_t)bytes, hipMemcpyHostToDevice, stream()));
  HIPCHK(hipStreamSynchronize(stream()));
}

void mpk_memcpy_dtoh(int64_t dst, int64_t src, int64_t bytes) {
  if (bytes <= 0 || !ctx()) return;
  HIPCHK(hipMemcpyAsync(reinterpret_cast<void *>(dst), reinterpret_cast<const void *>(src),
                        (size_t)bytes, hipMemcpyDeviceToHost, stream()));
  HIPCHK(hipStreamSynchronize(stream()));
}

void mpk_memcpy_dtod(int64_t dst, int64_t src, int64_t bytes) {
  if (bytes <= 0 || !ctx()) return;
  HIPCHK(hipMemcpyAsync(reinterpret_cast<void *>(dst), reinterpret_cast<const void *>(src),
                        (size_t)bytes, hipMemcpyDeviceToDevice, stream()));
}

void mpk_memset(int64_t dst, int64_t value, int64_t bytes) {
  if (bytes <= 0 || !ctx()) return;
  HIPCHK(hipMemsetAsync(reinterpret_cast<void *>(dst), (int)value, (size_t)bytes, stream()));
}

void mpk_synchronize(void) { sync(); }

int64_t mpk_set_reduction_order(int64_t order) {
  const int64_t old = red_ref() ? 1 : 0;
  if (order != 0 && order != 1) {
    set_error(4, "mpk_set_reduction_order: order %lld is neither 0 (tile tree) nor 1 (reference)",
              (long long)order);
    return old;
  }
  g_red_order.store((int)order, std::memory_order_relaxed);
  return old;
}

int64_t mpk_get_reduction_order(void) { return red_ref() ? 1 : 0; }

int64_t mpk_set_sweep_wait_limit(int64_t polls) {
  const int64_t old = g_watch_limit.load();
  if (polls < 1 || polls > 0xffffffffLL) {
    set_error(4, "mpk_set_sweep_wait_limit: %lld polls is outside 1..2^32-1", (long long)polls);
    return old;
  }
  g_watch_limit.store((unsigned)polls);
  g_watch_gen.fetch_add(1u);
  return old;
}

int64_t mpk_sweep_abort(int64_t on) {
  Ctx *c = ctx();
  if (!c || !c->watch) return -1;
  volatile unsigned *w = c->watch;
  w[0] = on ? 1u : 0u;
  return 0;
}

}  // extern "C"
