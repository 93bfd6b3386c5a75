// Byte-gather ceiling of the device: how many one-byte loads per second the
// CUs can issue against a working set of a given size, with nothing else in
// the loop.  The ray walk reads one uint8 label per voxel step; this is the
// rate it would reach if its stepping arithmetic cost nothing, measured at the
// volume's own size so the caches hold what they hold for the walk.
//
// Two access patterns:
//   ANACOR_GATHER_LANES: every lane of a wave reads its own pseudo-random
//                        byte (64 cache lines per wave-load, the worst case);
//   ANACOR_GATHER_ROW:   the 64 lanes read 64 consecutive bytes of one
//                        pseudo-random 64-byte row (one line per wave-load;
//                        the walk's waves, whose lanes run along x, touch
//                        about two lines per load).
// Addresses are a Weyl sequence (x += odd constant) masked to the power-of-two
// working set, two VALU operations per load, so the loop is bound by the
// memory pipeline and not by address arithmetic.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "anacor_mi355x.h"
#include "errors.h"

namespace anacor {
namespace {

constexpr int kProbeBlock = 256;
constexpr int kProbeUnroll = 8;       // loads in flight per lane
constexpr unsigned kWeyl = 0x9E3779B1u;

__device__ __forceinline__ unsigned mix32(unsigned x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}

// Each lane performs iters loads (iters a multiple of kProbeUnroll); offsets
// are masked to [0, mask], inside the buffer of mask + 1 bytes, and the
// buffer resource's range check would return 0 past its end in any case.
template <int PATTERN>
__global__ __launch_bounds__(kProbeBlock) void gather_probe_kernel(const uint8_t* buf,
                                                                   unsigned mask, int iters,
                                                                   unsigned* out) {
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)buf, (short)0, (int)(mask + 1u), 0x00020000);
  const unsigned gid = blockIdx.x * kProbeBlock + threadIdx.x;
  const unsigned lane = threadIdx.x & 63u;
  unsigned acc = 0;
  if (PATTERN == ANACOR_GATHER_LANES) {
    unsigned x = mix32(gid);
    for (int i = 0; i < iters; i += kProbeUnroll) {
#pragma unroll
      for (int u = 0; u < kProbeUnroll; ++u) {
        x += kWeyl;
        acc += __builtin_amdgcn_raw_buffer_load_b8(rsrc, (int)(x & mask), 0, 0);
      }
    }
  } else {
    // Wave-uniform row start in a scalar register, lane offset in the vector.
    unsigned x = __builtin_amdgcn_readfirstlane(mix32(gid >> 6));
    const unsigned row_mask = mask & ~63u;
    for (int i = 0; i < iters; i += kProbeUnroll) {
#pragma unroll
      for (int u = 0; u < kProbeUnroll; ++u) {
        x += kWeyl;
        acc += __builtin_amdgcn_raw_buffer_load_b8(rsrc, (int)lane, (int)(x & row_mask), 0);
      }
    }
  }
  out[gid] = acc;
}

}  // namespace
}  // namespace anacor

extern "C" int anacor_probe_gather(int device, int log2_bytes, int pattern, int64_t loads,
                                   double* loads_per_s) {
  using namespace anacor;
  clear_error();
  if (!loads_per_s || log2_bytes < 6 || log2_bytes > 30 ||
      (pattern != ANACOR_GATHER_LANES && pattern != ANACOR_GATHER_ROW) || loads <= 0) {
    set_error(ANACOR_ERR_ARG, "probe_gather: bad arguments (log2_bytes %d, pattern %d)",
              log2_bytes, pattern);
    return ANACOR_ERR_ARG;
  }
  *loads_per_s = 0.0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
    set_error(ANACOR_ERR_NO_DEVICE, "probe_gather: no HIP device %d", device);
    return ANACOR_ERR_NO_DEVICE;
  }
  // 8192 workgroups (32 per CU) of 256 lanes; iterations per lane from the
  // requested load count, rounded up to the unroll.
  const int blocks = 8192;
  const int64_t lanes = (int64_t)blocks * kProbeBlock;
  int64_t it = (loads + lanes - 1) / lanes;
  it = (it + kProbeUnroll - 1) / kProbeUnroll * kProbeUnroll;
  if (it > (1 << 15)) it = 1 << 15;
  const int iters = (int)it;
  const size_t nbytes = (size_t)1 << log2_bytes;
  const unsigned mask = (unsigned)(nbytes - 1);

  int rc = ANACOR_OK;
  uint8_t* buf = nullptr;
  unsigned* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t s = nullptr;
  auto check = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == ANACOR_OK) {
      set_error(ANACOR_ERR_HIP, "probe_gather: %s failed: %s", what, hipGetErrorString(e));
      rc = ANACOR_ERR_HIP;
    }
    return rc == ANACOR_OK;
  };
  if (check(hipSetDevice(device), "hipSetDevice") &&
      check(hipStreamCreate(&s), "hipStreamCreate") &&
      check(hipMalloc(&buf, nbytes), "hipMalloc") &&
      check(hipMalloc(&out, (size_t)lanes * sizeof(unsigned)), "hipMalloc") &&
      check(hipMemsetAsync(buf, 1, nbytes, s), "hipMemsetAsync") &&
      check(hipEventCreate(&e0), "hipEventCreate") && check(hipEventCreate(&e1), "hipEventCreate")) {
    auto launch = [&]() {
      if (pattern == ANACOR_GATHER_LANES)
        gather_probe_kernel<ANACOR_GATHER_LANES><<<blocks, kProbeBlock, 0, s>>>(buf, mask, iters, out);
      else
        gather_probe_kernel<ANACOR_GATHER_ROW><<<blocks, kProbeBlock, 0, s>>>(buf, mask, iters, out);
      return check(hipGetLastError(), "gather_probe_kernel launch");
    };
    const int reps = 5;
    bool ok = launch() && check(hipEventRecord(e0, s), "hipEventRecord");
    for (int r = 0; ok && r < reps; ++r) ok = launch();
    if (ok && check(hipEventRecord(e1, s), "hipEventRecord") &&
        check(hipEventSynchronize(e1), "hipEventSynchronize")) {
      float ms = 0.f;
      if (check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime") && ms > 0.f)
        *loads_per_s = (double)reps * (double)lanes * (double)iters / (ms * 1e-3);
    }
  }
  if (s) (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  if (buf) (void)hipFree(buf);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}
