/* Stand-in for the HIP runtime, test infrastructure only: it lets the library's
 * host code run on the CPU. Device memory comes from a Python-owned arena and
 * kernel launches are handed to a Python callback that interprets the
 * compiled gfx950 code (tests/isa/gcnemu.py). A launch whose interpretation
 * faults makes every later call report hipErrorIllegalAddress (700), as an
 * asynchronous fault would. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint32_t x, y, z; } dim3;
typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;

typedef uint64_t (*alloc_cb_t)(uint64_t nbytes);
typedef int (*free_cb_t)(uint64_t addr);
typedef int (*launch_cb_t)(const char* name, uint32_t gx, uint32_t gy, uint32_t gz, uint32_t bx, uint32_t by,
                           uint32_t bz, void** args, uint64_t shmem);

static alloc_cb_t g_alloc;
static free_cb_t g_free;
static launch_cb_t g_launch;
static hipError_t g_last = 0;
static hipError_t g_sticky = 0;

#define MAXF 4096
static const void* g_stub[MAXF];
static const char* g_name[MAXF];
static int g_nf;

void fakehip_set_callbacks(alloc_cb_t a, free_cb_t f, launch_cb_t l) { g_alloc = a; g_free = f; g_launch = l; }
int fakehip_sticky(void) { return g_sticky; }

void** __hipRegisterFatBinary(const void* data) { static void* h[4]; (void)data; return h; }
void __hipUnregisterFatBinary(void** h) { (void)h; }
void __hipRegisterVar(void** modules, void* var, char* hostVar, const char* deviceVar, int ext, size_t size, int constant,
                      int global) {
  (void)modules; (void)var; (void)hostVar; (void)deviceVar; (void)ext; (void)size; (void)constant; (void)global;
}
void __hipRegisterFunction(void** modules, const void* hostFunction, char* deviceFunction, const char* deviceName,
                           unsigned int threadLimit, void* tid, void* bid, dim3* blockDim, dim3* gridDim,
                           int* wSize) {
  (void)modules; (void)deviceFunction; (void)threadLimit; (void)tid; (void)bid; (void)blockDim; (void)gridDim;
  (void)wSize;
  if (g_nf < MAXF) { g_stub[g_nf] = hostFunction; g_name[g_nf] = deviceName; ++g_nf; }
}

static dim3 c_grid, c_block;
static size_t c_shmem;
static hipStream_t c_stream;
hipError_t __hipPushCallConfiguration(dim3 g, dim3 b, size_t sh, hipStream_t st) {
  c_grid = g; c_block = b; c_shmem = sh; c_stream = st; return 0;
}
hipError_t __hipPopCallConfiguration(dim3* g, dim3* b, size_t* sh, hipStream_t* st) {
  *g = c_grid; *b = c_block; *sh = c_shmem; *st = c_stream; return 0;
}
hipError_t hipLaunchKernel(const void* f, dim3 g, dim3 b, void** args, size_t sh, hipStream_t st) {
  (void)st;
  if (g_sticky) return g_sticky;
  const char* name = NULL;
  for (int i = 0; i < g_nf; ++i)
    if (g_stub[i] == f) { name = g_name[i]; break; }
  if (!name) { fprintf(stderr, "fakehip: unregistered kernel stub\n"); g_last = 98; return 98; }
  if (g.x == 0 || g.y == 0 || g.z == 0 || b.x == 0 || b.x * b.y * b.z > 1024) { g_last = 9; return 9; }
  int rc = g_launch(name, g.x, g.y, g.z, b.x, b.y, b.z, args, sh);
  if (rc) { g_sticky = 700; g_last = 700; return 0; } /* reported like an async fault */
  return 0;
}
hipError_t hipMalloc(void** p, size_t n) { *p = (void*)(uintptr_t)g_alloc(n); return *p ? 0 : 2; }
hipError_t hipFree(void* p) { if (p) g_free((uint64_t)(uintptr_t)p); return 0; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, int kind) { (void)kind; if (g_sticky) return g_sticky; memmove(d, s, n); return 0; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int kind, hipStream_t st) {
  (void)kind; (void)st; if (g_sticky) return g_sticky; memmove(d, s, n); return 0;
}
hipError_t hipMemcpy2D(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h, int kind) {
  (void)kind;
  if (g_sticky) return g_sticky;
  for (size_t r = 0; r < h; ++r) memmove((char*)d + r * dp, (const char*)s + r * sp, w);
  return 0;
}
hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h, int kind, hipStream_t st) {
  (void)st;
  return hipMemcpy2D(d, dp, s, sp, w, h, kind);
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st) { (void)st; memset(d, v, n); return 0; }
hipError_t hipStreamSynchronize(hipStream_t st) { (void)st; return g_sticky; }
hipError_t hipGetLastError(void) { hipError_t e = g_last; g_last = 0; return e; }
const char* hipGetErrorString(hipError_t e) { return e == 700 ? "an illegal memory access was encountered (interpreted)" : "fakehip error"; }
hipError_t hipGetDevice(int* d) { *d = 0; return 0; }
/* device symbols are not modelled: the allocation-free entry point is not interpreted */
hipError_t hipGetSymbolAddress(void** p, const void* sym) { (void)sym; *p = 0; return 801; }
/* no free-memory figure: optional stores (the adjoint's stage states) are not taken */
hipError_t hipMemGetInfo(size_t* fre, size_t* tot) { *fre = 0; *tot = 0; return 0; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = (void*)1; return 0; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags) { (void)flags; *e = (void*)1; return 0; }
/* launches run to completion in order, so a second stream is the first one */
hipError_t hipStreamCreateWithFlags(hipStream_t* st, unsigned flags) { (void)flags; *st = (void*)2; return 0; }
hipError_t hipStreamDestroy(hipStream_t st) { (void)st; return 0; }
hipError_t hipStreamWaitEvent(hipStream_t st, hipEvent_t e, unsigned flags) { (void)st; (void)e; (void)flags; return 0; }
hipError_t hipEventDestroy(hipEvent_t e) { (void)e; return 0; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st) { (void)e; (void)st; return 0; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) { (void)a; (void)b; *ms = 0.f; return 0; }
