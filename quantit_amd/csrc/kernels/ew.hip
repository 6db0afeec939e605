// Data movement, elementwise and reduction kernels (HBM-bound).  Work is
// split into (descriptor, chunk) tasks of kChunk elements so a launch always
// has many work-groups regardless of how uneven the blocks are.
#include "../kernels.hpp"

namespace qtt {
namespace kern {

__global__ __launch_bounds__(256) void k_permute(double* __restrict__ base, const PermDesc* __restrict__ descs,
                                                 const ChunkTask* __restrict__ tasks) {
  const ChunkTask t = tasks[blockIdx.x];
  const PermDesc d = descs[t.desc];
  const int64_t start = int64_t(t.chunk) * kChunk;
  const int64_t end = min(start + int64_t(kChunk), d.numel);
  const double* src = base + d.src_off;
  double* dst = base + d.dst_off;
  for (int64_t e = start + threadIdx.x; e < end; e += blockDim.x) {
    int64_t rem = e, off = 0;
    for (int k = d.rank - 1; k >= 0; --k) {
      const int64_t dim = d.dims[k];
      const int64_t q = rem / dim;
      off += (rem - q * dim) * d.sstride[k];
      rem = q;
    }
    dst[e] = src[off];
  }
}

void permute_copy(hipStream_t s, double* base, const PermDesc* descs, const ChunkTask* tasks, int ntasks) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_permute, dim3(ntasks), dim3(256), 0, s, base, descs, tasks);
}

__global__ __launch_bounds__(256) void k_copy2d(double* __restrict__ base, const Copy2D* __restrict__ descs,
                                                const ChunkTask* __restrict__ tasks) {
  const ChunkTask t = tasks[blockIdx.x];
  const Copy2D d = descs[t.desc];
  const int64_t numel = int64_t(d.rows) * d.cols;
  const int64_t start = int64_t(t.chunk) * kChunk;
  const int64_t end = min(start + int64_t(kChunk), numel);
  double* dst = base + d.dst_off;
  if (d.src_off < 0) {
    for (int64_t e = start + threadIdx.x; e < end; e += blockDim.x) {
      const int64_t r = e / d.cols, c = e - r * d.cols;
      dst[r * d.dst_rs + c * d.dst_cs] = d.alpha;
    }
    return;
  }
  const double* src = base + d.src_off;
  for (int64_t e = start + threadIdx.x; e < end; e += blockDim.x) {
    const int64_t r = e / d.cols, c = e - r * d.cols;
    dst[r * d.dst_rs + c * d.dst_cs] = d.alpha * src[r * d.src_rs + c * d.src_cs];
  }
}

// dst[i] = sum over s < parts of src[s * n + i] (fixed order): the reduction of
// split-K partial products.  desc = {dst_off, src_off, n, parts}.
__global__ __launch_bounds__(256) void k_sum_parts(double* __restrict__ base, const SumParts* __restrict__ descs,
                                                   const ChunkTask* __restrict__ tasks) {
  const ChunkTask t = tasks[blockIdx.x];
  const SumParts d = descs[t.desc];
  const int64_t i0 = int64_t(t.chunk) * kChunk;
  const int64_t i1 = i0 + kChunk < d.n ? i0 + kChunk : d.n;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    double v = 0.0;
    for (int p = 0; p < d.parts; ++p) v += base[d.src_off + p * d.n + i];
    base[d.dst_off + i] = v;
  }
}

void sum_parts(hipStream_t s, double* base, const SumParts* descs, const ChunkTask* tasks, int ntasks) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_sum_parts, dim3(ntasks), dim3(256), 0, s, base, descs, tasks);
}

void copy2d(hipStream_t s, double* base, const Copy2D* descs, const ChunkTask* tasks, int ntasks) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_copy2d, dim3(ntasks), dim3(256), 0, s, base, descs, tasks);
}

__global__ __launch_bounds__(256) void k_ew_binary(double* __restrict__ base, const EwDesc* __restrict__ descs,
                                                   const ChunkTask* __restrict__ tasks, int op, double min_abs) {
  const ChunkTask t = tasks[blockIdx.x];
  const EwDesc d = descs[t.desc];
  const int64_t start = int64_t(t.chunk) * kChunk;
  const int64_t end = min(start + int64_t(kChunk), d.numel);
  const double* x = base + d.x_off;
  const double* y = base + d.y_off;
  double* out = base + d.out_off;
  if (d.mode == 0) {
    const double yv = y[0];
    const bool skip = (op == 1) && (fabs(yv) < min_abs);
    for (int64_t e = start + threadIdx.x; e < end; e += blockDim.x)
      out[e] = skip ? x[e] : (op == 0 ? x[e] * yv : x[e] / yv);
  } else if (d.mode == 1) {
    for (int64_t e = start + threadIdx.x; e < end; e += blockDim.x) {
      const double yv = y[e % d.ncol];
      out[e] = op == 0 ? x[e] * yv : x[e] / yv;
    }
  } else {
    for (int64_t e = start + threadIdx.x; e < end; e += blockDim.x) out[e] = op == 0 ? x[e] * y[e] : x[e] / y[e];
  }
}

void ew_binary(hipStream_t s, double* base, const EwDesc* descs, const ChunkTask* tasks, int ntasks, int op,
               double min_abs) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_ew_binary, dim3(ntasks), dim3(256), 0, s, base, descs, tasks, op, min_abs);
}

__global__ __launch_bounds__(256) void k_axpby(double* __restrict__ base, const AxpbyDesc* __restrict__ descs,
                                               const ChunkTask* __restrict__ tasks, int64_t a_off, double a,
                                               int64_t b_off, double b) {
  const ChunkTask t = tasks[blockIdx.x];
  const AxpbyDesc d = descs[t.desc];
  const int64_t start = int64_t(t.chunk) * kChunk;
  const int64_t end = min(start + int64_t(kChunk), d.numel);
  const double av = a_off >= 0 ? base[a_off] : a;
  const double bv = b_off >= 0 ? base[b_off] : b;
  double* out = base + d.out_off;
  const double* x = d.x_off >= 0 ? base + d.x_off : nullptr;
  const double* y = d.y_off >= 0 ? base + d.y_off : nullptr;
  for (int64_t e = start + threadIdx.x; e < end; e += blockDim.x) {
    double v;
    if (x && y) {
      const double px = av * x[e];
      const double py = bv * y[e];
      v = px + py;
    } else if (x) {
      v = av * x[e];
    } else {
      v = bv * y[e];
    }
    out[e] = v;
  }
}

void axpby(hipStream_t s, double* base, const AxpbyDesc* descs, const ChunkTask* tasks, int ntasks, int64_t a_off,
           double a, int64_t b_off, double b) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3(ntasks), dim3(256), 0, s, base, descs, tasks, a_off, a, b_off, b);
}

// Block-wide sum in fixed order (deterministic).
__device__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void k_dot_partial(double* __restrict__ base, const DotDesc* __restrict__ descs,
                                                     const ChunkTask* __restrict__ tasks, int64_t partial_off) {
  __shared__ double sh[256];
  const ChunkTask t = tasks[blockIdx.x];
  const DotDesc d = descs[t.desc];
  const int64_t start = int64_t(t.chunk) * kChunk;
  const int64_t end = min(start + int64_t(kChunk), d.numel);
  const double* x = base + d.x_off;
  const double* y = base + d.y_off;
  double acc = 0.0;
  for (int64_t e = start + threadIdx.x; e < end; e += blockDim.x) acc += x[e] * y[e];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) base[partial_off + blockIdx.x] = r;
}

__global__ __launch_bounds__(256) void k_dot_final(double* __restrict__ base, int64_t partial_off, int n,
                                                   int64_t out_off, int sqrt_out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += base[partial_off + i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) base[out_off] = sqrt_out ? sqrt(r) : r;
}

void dot(hipStream_t s, double* base, const DotDesc* descs, const ChunkTask* tasks, int ntasks, int64_t partial_off,
         int64_t out_off, bool sqrt_out) {
  if (ntasks > 0)
    hipLaunchKernelGGL(k_dot_partial, dim3(ntasks), dim3(256), 0, s, base, descs, tasks, partial_off);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, base, partial_off, ntasks, out_off, int(sqrt_out));
}

__global__ void k_eig2x2(double* __restrict__ base, int64_t a0_off, int64_t a1_off, int64_t b_off, int64_t out_off,
                         int* flags) {
  if (threadIdx.x != 0) return;
  const double a0 = base[a0_off], a1 = base[a1_off], b = base[b_off];
  const double d = a0 - a1;
  const double crit = sqrt(d * d + 4.0 * (b * b));
  const double E0 = (a0 + a1 - crit) / 2.0;
  const double delt = E0 - a1;
  double o = sqrt(delt / (-crit));
  double n = (b * o) / delt;
  const bool zero_o = ((o + E0) == E0) || isnan(o);
  if (zero_o) {
    n = 1.0;
    o = 0.0;
  }
  if (isnan(o) || isnan(n)) flags[0] = 1;
  base[out_off + 0] = E0;
  base[out_off + 1] = o;
  base[out_off + 2] = n;
}

void eig2x2(hipStream_t s, double* base, int64_t a0_off, int64_t a1_off, int64_t b_off, int64_t out_off,
            int* flags) {
  hipLaunchKernelGGL(k_eig2x2, dim3(1), dim3(64), 0, s, base, a0_off, a1_off, b_off, out_off, flags);
}

__global__ __launch_bounds__(256) void k_normalize(double* __restrict__ base, int64_t off, int64_t n) {
  __shared__ double sh[256];
  double* d = base + off;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += d[i] * d[i];
  const double nrm = sqrt(block_sum(acc, sh));
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) d[i] = d[i] / nrm;
}

void normalize(hipStream_t s, double* base, int64_t off, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_normalize, dim3(1), dim3(256), 0, s, base, off, n);
}

}  // namespace kern
}  // namespace qtt
