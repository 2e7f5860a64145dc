// tomo_cvss.hip -- continuous support volume and its gradient for gfx950.
//
// Vss_c(yaw, pitch) = sum over triangles t of sigma(-r.n_t - s) * A_t * (-r.n_t) * (r.c_t - zlow)
// with r = (-sin pitch, cos pitch sin yaw, cos pitch cos yaw) the build direction in mesh
// coordinates (the third row of Rz(roll) Ry(pitch) Rx(yaw), SMatrix4f.cpp:213-240, so roll drops
// out), s = sin(theta_c), n_t the unit face normal, A_t the area and c_t the centroid.  It is the
// projected-area times height form of the support volume whose per-triangle precedent is
// getCHVss (STomoNV_CvxH.cpp:97-174), measured from the build plate instead of the convex hull.
// sigma is a C1 smoothstep of width tau around the critical angle (tau = 0: the hard step) and
// zlow = -T log sum_v exp(-r.v / T) a soft minimum of the vertex heights (T = 0: the lowest
// vertex): a flat triangle otherwise switches on as a whole, and the lowest vertex changes
// discretely, which would leave jumps and kinks in Vss_c.
//
// Forward, per orientation: the lowest vertex (one 64-bit atomicMin per work-group over
// (ordered float bits, index), ties to the lowest index); with T > 0 the softmin weights sum
// e_v = exp((zmin - r.v) / T) and sum e_v v; then the triangle sum.  Terms are fp32, sums fp64
// with one atomic per work-group.  Backward: d Vss_c / d r = sum_t [dw_t/dd n_t (r.c_t - zlow)
// + w_t (c_t - v*)], w_t = sigma A_t (-d), d = r.n_t, v* = sum e_v v / sum e_v (the lowest vertex
// when T = 0), then the chain rule through r(yaw, pitch).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tomo {

namespace {

constexpr int kCvssThreads = 256;
constexpr int kCvssPerThread = 8;  // elements per thread per work-group
constexpr int kCvssChunk = kCvssThreads * kCvssPerThread;
constexpr int kAux = 12;  // doubles per orientation in the workspace
enum { AUX_ZKEY = 0, AUX_S = 1, AUX_E = 2, AUX_EV = 3, AUX_G = 6 };

struct CvssArgs {
  // mesh, structure of arrays, coordinates relative to the mesh's AABB centre
  const float *nx, *ny, *nz, *area, *cx, *cy, *cz;  // [n_tri]
  const float *vx, *vy, *vz;                        // [n_vtx]
  int32_t n_tri, n_vtx;
  float s;                                          // sin(theta_c)
  float tau, temp;                                  // smoothstep width, softmin temperature
  // batch
  int32_t n;
  const float *ypr;      // [n][3] radians
  double *aux;           // [n][kAux]: zkey (uint64 bits), S, sum e, sum e v (3), g (3)
  float *vss;            // [n]
  const float *grad_vss; // [n]
  float *grad_ypr;       // [n][3]
  int32_t tri_blocks, vtx_blocks;
};

__device__ __forceinline__ void build_dir(const float *ypr, float &r0, float &r1, float &r2) {
  const float sy = sinf(ypr[0]), cy = cosf(ypr[0]);
  const float sp = sinf(ypr[1]), cp = cosf(ypr[1]);
  r0 = -sp;
  r1 = cp * sy;
  r2 = cp * cy;
}

__device__ __forceinline__ uint32_t ordered_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

template <typename T>
__device__ __forceinline__ T wave_red_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum N doubles over the work-group; thread 0 adds them to dst[0..N).
template <int N>
__device__ __forceinline__ void block_add(double (&v)[N], double *dst) {
  __shared__ double red[kCvssThreads / 64][N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; k++) v[k] = wave_red_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; k++) red[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      double t = 0.0;
      for (int w = 0; w < kCvssThreads / 64; w++) t += red[w][k];
      if (t != 0.0) atomicAdd(dst + k, t);
    }
  }
}

__global__ __launch_bounds__(kCvssThreads) void cvss_init_kernel(CvssArgs a, int backward) {
  const int32_t b = blockIdx.x * kCvssThreads + threadIdx.x;
  if (b >= a.n) return;
  double *x = a.aux + (int64_t)b * kAux;
  if (!backward) {
    reinterpret_cast<unsigned long long *>(x)[AUX_ZKEY] = ~0ull;
    x[AUX_S] = 0.0;
    x[AUX_E] = 0.0;
    x[AUX_EV] = x[AUX_EV + 1] = x[AUX_EV + 2] = 0.0;
  }
  x[AUX_G] = x[AUX_G + 1] = x[AUX_G + 2] = 0.0;
}

// Lowest vertex along r: key = ordered bits of r.v (high word) | vertex index (low word).
__global__ __launch_bounds__(kCvssThreads) void cvss_zmin_kernel(CvssArgs a) {
  const int32_t b = blockIdx.x / a.vtx_blocks, chunk = blockIdx.x % a.vtx_blocks;
  if (b >= a.n) return;
  float r0, r1, r2;
  build_dir(a.ypr + (int64_t)b * 3, r0, r1, r2);
  unsigned long long best = ~0ull;
  const int32_t v0 = chunk * kCvssChunk, v1 = min(v0 + kCvssChunk, a.n_vtx);
  for (int32_t v = v0 + threadIdx.x; v < v1; v += kCvssThreads) {
    const float z = r0 * a.vx[v] + r1 * a.vy[v] + r2 * a.vz[v];
    const unsigned long long k = ((unsigned long long)ordered_bits(z) << 32) | (uint32_t)v;
    best = k < best ? k : best;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long q = __shfl_xor(best, o);
    best = q < best ? q : best;
  }
  __shared__ unsigned long long red[kCvssThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kCvssThreads / 64; w++) best = red[w] < best ? red[w] : best;
    atomicMin(reinterpret_cast<unsigned long long *>(a.aux + (int64_t)b * kAux) + AUX_ZKEY, best);
  }
}

__device__ __forceinline__ float hard_low(const CvssArgs &a, int32_t b, float r0, float r1, float r2,
                                          int32_t &vi) {
  const unsigned long long k =
      reinterpret_cast<const unsigned long long *>(a.aux + (int64_t)b * kAux)[AUX_ZKEY];
  vi = (int32_t)(uint32_t)(k & 0xffffffffu);
  return r0 * a.vx[vi] + r1 * a.vy[vi] + r2 * a.vz[vi];  // same value the key was made from
}

// Softmin weights: sum e_v and sum e_v v over vertices within 40 T of the lowest one (the rest
// contribute below exp(-40)).
__global__ __launch_bounds__(kCvssThreads) void cvss_soft_kernel(CvssArgs a) {
  const int32_t b = blockIdx.x / a.vtx_blocks, chunk = blockIdx.x % a.vtx_blocks;
  if (b >= a.n) return;
  float r0, r1, r2;
  build_dir(a.ypr + (int64_t)b * 3, r0, r1, r2);
  int32_t vi;
  const float zmin = hard_low(a, b, r0, r1, r2, vi);
  const float cut = 40.0f * a.temp;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int32_t v0 = chunk * kCvssChunk, v1 = min(v0 + kCvssChunk, a.n_vtx);
  for (int32_t v = v0 + threadIdx.x; v < v1; v += kCvssThreads) {
    const float x = a.vx[v], y = a.vy[v], z = a.vz[v];
    const float dz = (r0 * x + r1 * y + r2 * z) - zmin;
    if (dz < cut) {
      const double e = exp(-(double)dz / (double)a.temp);
      acc[0] += e;
      acc[1] += e * (double)x;
      acc[2] += e * (double)y;
      acc[3] += e * (double)z;
    }
  }
  block_add<4>(acc, a.aux + (int64_t)b * kAux + AUX_E);
}

// zlow and v* = d zlow / d r for orientation b.
__device__ __forceinline__ float low_point(const CvssArgs &a, int32_t b, float r0, float r1, float r2,
                                           float &lx, float &ly, float &lz) {
  int32_t vi;
  const float zmin = hard_low(a, b, r0, r1, r2, vi);
  if (a.temp > 0.0f) {
    const double *x = a.aux + (int64_t)b * kAux;
    const double se = x[AUX_E];
    lx = (float)(x[AUX_EV] / se);
    ly = (float)(x[AUX_EV + 1] / se);
    lz = (float)(x[AUX_EV + 2] / se);
    return (float)((double)zmin - (double)a.temp * log(se));
  }
  lx = a.vx[vi];
  ly = a.vy[vi];
  lz = a.vz[vi];
  return zmin;
}

// Smoothstep sigma(x) of width tau and its derivative (tau = 0: the hard step x > 0).
__device__ __forceinline__ void smooth_step(float x, float tau, float &sig, float &dsig) {
  if (tau > 0.0f) {
    const float u = fminf(fmaxf(x / tau + 0.5f, 0.0f), 1.0f);
    sig = u * u * (3.0f - 2.0f * u);
    dsig = 6.0f * u * (1.0f - u) / tau;
  } else {
    sig = x > 0.0f ? 1.0f : 0.0f;
    dsig = 0.0f;
  }
}

__global__ __launch_bounds__(kCvssThreads) void cvss_sum_kernel(CvssArgs a) {
  const int32_t b = blockIdx.x / a.tri_blocks, chunk = blockIdx.x % a.tri_blocks;
  if (b >= a.n) return;
  float r0, r1, r2, lx, ly, lz;
  build_dir(a.ypr + (int64_t)b * 3, r0, r1, r2);
  const float zlow = low_point(a, b, r0, r1, r2, lx, ly, lz);
  const float lo = -0.5f * a.tau;  // sigma = 0 below
  double acc[1] = {0.0};
  const int32_t t0 = chunk * kCvssChunk, t1 = min(t0 + kCvssChunk, a.n_tri);
  for (int32_t t = t0 + threadIdx.x; t < t1; t += kCvssThreads) {
    const float d = r0 * a.nx[t] + r1 * a.ny[t] + r2 * a.nz[t];
    const float x = -d - a.s;
    if (x > lo) {
      float sig, dsig;
      smooth_step(x, a.tau, sig, dsig);
      const float h = (r0 * a.cx[t] + r1 * a.cy[t] + r2 * a.cz[t]) - zlow;
      acc[0] += (double)(sig * a.area[t] * (-d) * h);
    }
  }
  block_add<1>(acc, a.aux + (int64_t)b * kAux + AUX_S);
}

__global__ __launch_bounds__(kCvssThreads) void cvss_out_kernel(CvssArgs a) {
  const int32_t b = blockIdx.x * kCvssThreads + threadIdx.x;
  if (b >= a.n) return;
  a.vss[b] = (float)a.aux[(int64_t)b * kAux + AUX_S];
}

__global__ __launch_bounds__(kCvssThreads) void cvss_grad_kernel(CvssArgs a) {
  const int32_t b = blockIdx.x / a.tri_blocks, chunk = blockIdx.x % a.tri_blocks;
  if (b >= a.n) return;
  float r0, r1, r2, lx, ly, lz;
  build_dir(a.ypr + (int64_t)b * 3, r0, r1, r2);
  const float zlow = low_point(a, b, r0, r1, r2, lx, ly, lz);
  const float lo = -0.5f * a.tau;
  double g[3] = {0.0, 0.0, 0.0};
  const int32_t t0 = chunk * kCvssChunk, t1 = min(t0 + kCvssChunk, a.n_tri);
  for (int32_t t = t0 + threadIdx.x; t < t1; t += kCvssThreads) {
    const float nx = a.nx[t], ny = a.ny[t], nz = a.nz[t];
    const float d = r0 * nx + r1 * ny + r2 * nz;
    const float x = -d - a.s;
    if (x > lo) {
      float sig, dsig;
      smooth_step(x, a.tau, sig, dsig);
      const float A = a.area[t];
      const float cx = a.cx[t], cy = a.cy[t], cz = a.cz[t];
      const float h = (r0 * cx + r1 * cy + r2 * cz) - zlow;
      const float w = sig * A * (-d);
      const float dh = -A * (sig + (-d) * dsig) * h;  // dw/dd * h
      g[0] += (double)(dh * nx + w * (cx - lx));
      g[1] += (double)(dh * ny + w * (cy - ly));
      g[2] += (double)(dh * nz + w * (cz - lz));
    }
  }
  block_add<3>(g, a.aux + (int64_t)b * kAux + AUX_G);
}

// d r / d yaw = (0, cp cy, -cp sy); d r / d pitch = (-cp, -sp sy, -sp cy); d r / d roll = 0.
__global__ __launch_bounds__(kCvssThreads) void cvss_chain_kernel(CvssArgs a) {
  const int32_t b = blockIdx.x * kCvssThreads + threadIdx.x;
  if (b >= a.n) return;
  const float *y = a.ypr + (int64_t)b * 3;
  const double sy = sin((double)y[0]), cy = cos((double)y[0]);
  const double sp = sin((double)y[1]), cp = cos((double)y[1]);
  const double *g = a.aux + (int64_t)b * kAux + AUX_G;
  const double go = (double)a.grad_vss[b];
  const double dyaw = g[1] * (cp * cy) + g[2] * (-cp * sy);
  const double dpitch = g[0] * (-cp) + g[1] * (-sp * sy) + g[2] * (-sp * cy);
  a.grad_ypr[(int64_t)b * 3 + 0] = (float)(go * dyaw);
  a.grad_ypr[(int64_t)b * 3 + 1] = (float)(go * dpitch);
  a.grad_ypr[(int64_t)b * 3 + 2] = 0.0f;
}

}  // namespace

// Mesh arrays as laid out in device memory by the host (structure of arrays).
struct CvssMeshDev {
  const float *nx, *ny, *nz, *area, *cx, *cy, *cz, *vx, *vy, *vz;
  int32_t n_tri, n_vtx;
  float s, tau, temp;
};

static CvssArgs make_args(const CvssMeshDev &m, int32_t n, const float *ypr, double *aux) {
  CvssArgs a{};
  a.nx = m.nx; a.ny = m.ny; a.nz = m.nz; a.area = m.area; a.cx = m.cx; a.cy = m.cy; a.cz = m.cz;
  a.vx = m.vx; a.vy = m.vy; a.vz = m.vz;
  a.n_tri = m.n_tri; a.n_vtx = m.n_vtx; a.s = m.s; a.tau = m.tau; a.temp = m.temp;
  a.n = n; a.ypr = ypr; a.aux = aux;
  a.tri_blocks = (m.n_tri + kCvssChunk - 1) / kCvssChunk;
  a.vtx_blocks = (m.n_vtx + kCvssChunk - 1) / kCvssChunk;
  if (a.tri_blocks < 1) a.tri_blocks = 1;
  if (a.vtx_blocks < 1) a.vtx_blocks = 1;
  return a;
}

hipError_t cvss_forward(const CvssMeshDev &m, int32_t n, const float *ypr, float *vss, double *aux,
                        hipStream_t st) {
  if (n <= 0) return hipSuccess;
  CvssArgs a = make_args(m, n, ypr, aux);
  a.vss = vss;
  const int nb = (n + kCvssThreads - 1) / kCvssThreads;
  hipLaunchKernelGGL(cvss_init_kernel, dim3(nb), dim3(kCvssThreads), 0, st, a, 0);
  if (m.n_vtx > 0)
    hipLaunchKernelGGL(cvss_zmin_kernel, dim3((int64_t)n * a.vtx_blocks), dim3(kCvssThreads), 0, st, a);
  if (m.n_vtx > 0 && m.temp > 0.0f)
    hipLaunchKernelGGL(cvss_soft_kernel, dim3((int64_t)n * a.vtx_blocks), dim3(kCvssThreads), 0, st, a);
  if (m.n_tri > 0 && m.n_vtx > 0)
    hipLaunchKernelGGL(cvss_sum_kernel, dim3((int64_t)n * a.tri_blocks), dim3(kCvssThreads), 0, st, a);
  hipLaunchKernelGGL(cvss_out_kernel, dim3(nb), dim3(kCvssThreads), 0, st, a);
  return hipGetLastError();
}

hipError_t cvss_backward(const CvssMeshDev &m, int32_t n, const float *ypr, double *aux,
                         const float *grad_vss, float *grad_ypr, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  CvssArgs a = make_args(m, n, ypr, aux);
  a.grad_vss = grad_vss;
  a.grad_ypr = grad_ypr;
  const int nb = (n + kCvssThreads - 1) / kCvssThreads;
  hipLaunchKernelGGL(cvss_init_kernel, dim3(nb), dim3(kCvssThreads), 0, st, a, 1);
  if (m.n_tri > 0 && m.n_vtx > 0)
    hipLaunchKernelGGL(cvss_grad_kernel, dim3((int64_t)n * a.tri_blocks), dim3(kCvssThreads), 0, st, a);
  hipLaunchKernelGGL(cvss_chain_kernel, dim3(nb), dim3(kCvssThreads), 0, st, a);
  return hipGetLastError();
}

}  // namespace tomo
