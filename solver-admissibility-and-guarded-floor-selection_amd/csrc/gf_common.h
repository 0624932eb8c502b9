// Shared definitions for the gfhip library (gfx950 / CDNA4, wave64).
//
// Layout conventions (identical at every multigrid level):
//   node (ix,iy,iz) of an (nx,ny,nz)-element grid has linear index
//   n = (ix*(ny+1) + iy)*(nz+1) + iz, DOF = 3n + component.
//   Vectors are indexed by *free* DOF (the reference's u_free layout); a
//   DofMap turns a full DOF into its free index or -1.
//   Coarse operators are 27-point 3x3-block stencils stored SoA:
//   A[(k*9 + r*3 + c)*n_nodes + n], k = (dx+1)*9 + (dy+1)*3 + (dz+1).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

namespace gf {

void set_error(const std::string& msg);
const char* last_error();

#define GF_HIP(call)                                                          \
  do {                                                                        \
    hipError_t _e = (call);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::gf::set_error(std::string(#call) + ": " + hipGetErrorString(_e));     \
      return -1;                                                              \
    }                                                                         \
  } while (0)

#define GF_CHECK_LAUNCH()                                                     \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess) {                                                   \
      ::gf::set_error(std::string("kernel launch: ") + hipGetErrorString(_e)); \
      return -1;                                                              \
    }                                                                         \
  } while (0)

#define GF_TRY(expr)                                                          \
  do {                                                                        \
    int _rc = (expr);                                                         \
    if (_rc != 0) return _rc;                                                 \
  } while (0)

constexpr int kBlock = 256;
constexpr int kRedBlocks = 512;  // fixed reduction grid => deterministic sums

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__host__ __device__ inline unsigned short bf16_bits_rne(float f) {
  unsigned int b;
  __builtin_memcpy(&b, &f, 4);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (unsigned short)(b >> 16);
}

// Geometry of one level's node grid.
struct Grid {
  int nx, ny, nz;      // elements
  int NX, NY, NZ;      // nodes = elements + 1
  int n_nodes;
  int ndof;
  int shape;           // fine-operator launch shape (0 = process default)
  int shape32;         // its FP32 operator's shape (0 = as shape)
};

inline Grid make_grid(int nx, int ny, int nz) {
  Grid g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.NX = nx + 1; g.NY = ny + 1; g.NZ = nz + 1;
  g.n_nodes = g.NX * g.NY * g.NZ;
  g.ndof = 3 * g.n_nodes;
  g.shape = 0;
  g.shape32 = 0;
  return g;
}

// Full DOF -> free index.  Offset mode (off >= 0): the free DOFs are exactly
// [off, ndof) -- true for any face clamp at x = 0 -- so the index is computed,
// not loaded.  Table mode (off < 0): f2f[dof].
struct DofMap {
  int off;
  const int* f2f;
  __device__ __forceinline__ int operator()(int dof) const {
    if (off >= 0) return dof >= off ? dof - off : -1;
    return f2f[dof];
  }
};

// local node index of position (px,py,pz) in the reference hex ordering
// (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1)
__host__ __device__ constexpr int lidx(int px, int py, int pz) {
  return pz * 4 + (py ? (px ? 2 : 3) : (px ? 1 : 0));
}
__host__ __device__ constexpr int lpos(int j, int axis) {
  // inverse of lidx
  return axis == 2 ? (j >> 2)
                   : (axis == 1 ? (((j & 3) == 2 || (j & 3) == 3) ? 1 : 0)
                                : (((j & 3) == 1 || (j & 3) == 2) ? 1 : 0));
}

// Workgroups are dispatched round-robin over the 8 XCDs of MI355X (block b
// runs on XCD b % 8) and each XCD has its own L2.  xcd_tile(b, nb) renumbers
// blocks so that the blocks of one XCD take one contiguous range of tiles:
// neighbouring tiles then share an L2 and a stencil's halo is fetched from
// HBM once per XCD range instead of once per tile.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int kXcd = 8;
  const int q = nb / kXcd, r = nb % kXcd, k = b % kXcd;
  return k * q + (k < r ? k : r) + b / kXcd;
}

// Serialises device allocation, freeing and legacy-stream set-up calls with
// hipGraph capture across host threads: a hipMalloc/hipFree/hipMemset issued
// by one thread while another thread's stream is capturing can fail (or
// invalidate the capture), so every such site and every capture holds this
// lock.  Recursive: release paths run inside allocation paths.
std::recursive_mutex& device_mutex();
#define GF_DEVICE_LOCK() std::lock_guard<std::recursive_mutex> _gf_dev_lock(::gf::device_mutex())

// Set-up writes of device memory that later work on any stream reads (host
// tables, zeroed flags and scalars).  They run on a private non-blocking
// stream and are waited for before returning, so they are complete before
// the caller enqueues anything: a null-stream hipMemset/hipMemcpy is not
// ordered with the non-blocking streams the solvers run on, and a late one
// can land in the middle of a solve.
int upload_now(void* dst, const void* src, size_t bytes);
int zero_now(void* dst, size_t bytes);

// Wait until every memory operation this lane issued has been acknowledged
// (s_waitcnt with all counters at zero; on gfx9 stores count in vmcnt).  The
// signal fences keep the compiler from moving memory accesses across it.
__device__ __forceinline__ void store_ack_wait() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Free index of a full DOF with the map mode fixed at compile time.
template <bool OFF>
__device__ __forceinline__ int map_idx(const DofMap& m, int dof) {
  if (OFF) return dof >= m.off ? dof - m.off : -1;
  return m.f2f[dof];
}

// Branch-free read of u at full DOF `dof` through the map: DOFs outside the
// grid (in == false) or fixed read as 0.  The load always targets a valid
// address (index 0 when masked), so every load of a window can be issued
// back to back instead of behind an exec-mask branch.
// The index is taken unsigned (a free index is below 2^31), so the load can
// use a scalar base with a 32-bit lane offset and needs no sign extension.
template <bool OFF, typename T>
__device__ __forceinline__ T map_load(const DofMap& m, const T* __restrict__ u, int dof, bool in) {
  int f;
  if (OFF) {
    f = dof - m.off;
  } else {
    f = m.f2f[in ? (unsigned)dof : 0u];
  }
  const bool ok = in & (f >= 0);
  const T v = u[ok ? (unsigned)f : 0u];
  return ok ? v : T(0);
}

// Launch gate of a kernel in a device-controlled sequence: a flag (nullable)
// written by an earlier kernel on the stream, so its value is uniform over
// the grid and visible from the kernel's start like any input.  A kernel
// requests it at its top and tests it just before its first global store
// (or grid-wide exchange), so the flag's load latency passes under the
// kernel's own loads; a closed gate then costs only loads of the launch's
// valid buffers.  The flag is read by a vector load (a relaxed atomic load
// is never a scalar-cache load): vector loads complete in order, so the
// test waits for the flag only, while a scalar load would be waited for at
// the first use of any later scalar load (the kernel's arguments).
// A kernel without a gate reads this always-open flag instead, so the load
// is unconditional (a branch around it would be joined, and so waited for,
// at once).
static __device__ const int kGateAlwaysOpen = 1;

__device__ __forceinline__ bool gate_open(const int* gate) {
  // a global-memory load: a flat load would also count as a scalar-memory
  // access and be waited for with the kernel's arguments
  using GInt = const __attribute__((address_space(1))) int;
  GInt* p = (GInt*)(gate ? gate : &kGateAlwaysOpen);
  // a zero offset held in a vector register: the flag's address then looks
  // lane-varying to the compiler, which keeps the flag in a vector register
  // until its test (a uniform flag would be moved to a scalar register, and
  // waited for, right after its load)
  int zero;
  asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
  return __hip_atomic_load(p + zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
}

// Device scalar slots (FP64) used by device-resident control flow.
struct Scalars {
  double* s;   // base pointer of a device array of doubles
  int* ticket; // reduction tickets (one per concurrent reduction site)
};

}  // namespace gf
