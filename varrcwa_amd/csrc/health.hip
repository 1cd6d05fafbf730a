// Numerical-failure reporting. Each context owns a small device record that the
// factorizations and eigensolvers write when a matrix fails (a zero pivot, an
// unconverged QR iteration, a zero propagation constant); the top-level calls reset it
// on entry and read it back once on exit, so a failure surfaces as a positive status
// with a message instead of silent NaN/Inf results. Batched stack calls also flag each
// instance whose eigensolves failed or whose S-matrix holds non-finite values.
#include <cstdio>

#include "vrc_ops.hpp"

namespace vrc {

// h[0]: first failure code (0 none); h[1]: index of the failing matrix in its launch (or
// the flat index of a zero gamma); h[2]: the LAPACK-style info value. The first writer wins.
__global__ void health_info_kernel(const int* info, int batch, int code, int* h) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < batch; b += gridDim.x * blockDim.x) {
    const int v = info[b];
    if (v != 0 && atomicCAS(h, 0, code) == 0) {
      h[1] = b;
      h[2] = v;
    }
  }
}

__global__ void health_gamma_kernel(const cd* gam, long count, int* h) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
    const cd g = gam[i];
    if (g.x == 0.0 && g.y == 0.0 && atomicCAS(h, 0, (int)H_GAMMA0) == 0) {
      h[1] = (int)i;
      h[2] = 0;
    }
  }
}

// flags[inst] = code when instance inst's per-instance entries (per complex values at
// stride `stride`) hold a NaN or an infinity, unless a code is already set.
__global__ void nonfinite_kernel(const cd* A, long per, long stride, int* flags, int code) {
  const long inst = blockIdx.y;
  const cd* a = A + inst * stride;
  bool bad = false;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < per; i += (long)gridDim.x * blockDim.x) {
    const cd v = a[i];
    bad |= !(isfinite(v.x) && isfinite(v.y));
  }
  if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicCAS(flags + inst, 0, code);
}

// flags[base + owner[q]] = code for every job q whose info is non-zero.
__global__ void info_owner_kernel(const int* info, const int* owner, int njobs, int code, int* flags) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < njobs && info[q] != 0) atomicCAS(flags + owner[q], 0, code);
}

static int* active_health() {
  vrc_ctx* c = tl_ctx();
  return c ? c->health : nullptr;
}

void health_note_info(hipStream_t s, const int* info, int batch, int code) {
  int* h = active_health();
  if (!h || !info || batch <= 0) return;
  health_info_kernel<<<cdiv_up(batch, 256) < 64 ? cdiv_up(batch, 256) : 64, 256, 0, s>>>(info, batch, code, h);
  launch_check("health_info");
}

void health_note_gamma(hipStream_t s, const cd* gam, long count) {
  int* h = active_health();
  if (!h || !gam || count <= 0) return;
  const long blocks = (count + 255) / 256;
  health_gamma_kernel<<<(int)(blocks < 1024 ? blocks : 1024), 256, 0, s>>>(gam, count, h);
  launch_check("health_gamma");
}

void nonfinite_flags(hipStream_t s, const cd* A, long per, long stride, int ninst, int* flags, int code) {
  if (ninst <= 0 || per <= 0) return;
  const long blocks = (per + 255) / 256;
  dim3 g((unsigned)(blocks < 256 ? blocks : 256), ninst);
  nonfinite_kernel<<<g, 256, 0, s>>>(A, per, stride, flags, code);
  launch_check("nonfinite");
}

void info_to_owner(hipStream_t s, const int* info, const int* owner, int njobs, int code, int* flags) {
  if (njobs <= 0) return;
  info_owner_kernel<<<cdiv_up(njobs, 256), 256, 0, s>>>(info, owner, njobs, code, flags);
  launch_check("info_owner");
}

void health_reset(vrc_ctx& c) {
  c.host_code = 0;
  c.host_msg.clear();
  if (c.health) fill_int(c.stream, c.health, 0, HEALTH_INTS);
}

const char* health_text(int code) {
  switch (code) {
    case H_EIG: return "the eigensolver did not converge";
    case H_LU: return "exactly singular matrix (zero pivot) in an LU factorization";
    case H_GAMMA0: return "zero propagation constant gamma (Rayleigh/Wood anomaly: V = Q W / gamma is undefined)";
    case H_NONFINITE: return "non-finite values (NaN or infinity) in the results";
    case H_DEPTH: return "varRCWA subdivision reached its depth limit with the segment error above pe";
    default: return "numerical failure";
  }
}

int health_read(vrc_ctx& c, std::string& msg) {
  int h[HEALTH_INTS] = {0, 0, 0, 0};
  if (c.health) {
    VRC_HIP(hipMemcpyAsync(h, c.health, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    VRC_HIP(hipStreamSynchronize(c.stream));
  }
  if (h[0] != 0) {
    char buf[96];
    snprintf(buf, sizeof(buf), " (matrix %d of its launch, info %d)", h[1], h[2]);
    msg = std::string(health_text(h[0])) + buf;
    if (!c.host_msg.empty()) msg = c.host_msg + "; " + msg;
    return c.host_code ? c.host_code : h[0];
  }
  if (c.host_code) {
    msg = c.host_msg;
    return c.host_code;
  }
  return 0;
}

}  // namespace vrc
