// Multi-center Taylor reconstruction of a Fourier field on Yee points.
//
// u(r) = sum_j c_j exp(i q_j . r)
//      ~ sum_{|alpha|<=p} (r - r_c)^alpha * m_alpha(r_c),
// m_alpha(r_c) = i^|alpha|/alpha! * sum_j q_j^alpha exp(i q_j . r_c) c_j
// (taylor_on_blocks_exp3.m:4-6; reconstruct_component_taylor_batch_exp4.m:14-38).
// Term order alpha = (a1, a2, a3) with a1 outermost (build_multi_index_list_3d_exp3.m:6-12).
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "krylov_kernels.hpp"

namespace fame6d {

constexpr int kTaylorMaxP = 16;
constexpr int kTaylorChunk = 32;  // terms per workgroup in the moment kernel

__host__ __device__ inline int taylor_terms(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

// Moments: workgroup (c, chunk) accumulates terms [t0, t0+32) for center c.
__global__ __launch_bounds__(256) void k_taylor_moments(long long NF, const double* qx, const double* qy,
                                                        const double* qz, const cpx<double>* coeff, int p,
                                                        const double* centers, const int* alist,
                                                        const cpx<double>* calpha, int L,
                                                        cpx<double>* mom) {
  __shared__ double red[256];
  __shared__ int al[kTaylorChunk * 3];
  const int c = blockIdx.x;
  const int t0 = blockIdx.y * kTaylorChunk;
  const int tc = min(kTaylorChunk, L - t0);
  for (int i = threadIdx.x; i < tc * 3; i += blockDim.x) al[i] = alist[t0 * 3 + i];
  __syncthreads();
  const double rx = centers[3 * c], ry = centers[3 * c + 1], rz = centers[3 * c + 2];
  double ar[kTaylorChunk], ai[kTaylorChunk];
#pragma unroll
  for (int k = 0; k < kTaylorChunk; ++k) ar[k] = ai[k] = 0.0;
  for (long long j = blockIdx.z * (long long)blockDim.x + threadIdx.x; j < NF;
       j += (long long)gridDim.z * blockDim.x) {
    const double a = qx[j], b = qy[j], d = qz[j];
    double sn, cs;
    sincos(a * rx + b * ry + d * rz, &sn, &cs);
    const cpx<double> z = mk(cs, sn) * coeff[j];
    double px[kTaylorMaxP + 1], py[kTaylorMaxP + 1], pz[kTaylorMaxP + 1];
    px[0] = py[0] = pz[0] = 1.0;
    for (int s = 1; s <= p; ++s) {
      px[s] = px[s - 1] * a;
      py[s] = py[s - 1] * b;
      pz[s] = pz[s - 1] * d;
    }
#pragma unroll
    for (int k = 0; k < kTaylorChunk; ++k) {
      if (k < tc) {
        const double w = px[al[3 * k]] * py[al[3 * k + 1]] * pz[al[3 * k + 2]];
        ar[k] += w * z.re;
        ai[k] += w * z.im;
      }
    }
  }
  for (int k = 0; k < tc; ++k) {
    const double sr = block_sum<256>(ar[k], red);
    __syncthreads();
    const double si = block_sum<256>(ai[k], red);
    __syncthreads();
    if (threadIdx.x == 0) {
      // partial over grid.z slices accumulated by atomic-free second stage
      mom[((long long)blockIdx.z * gridDim.x + c) * L + t0 + k] = mk(sr, si);
    }
  }
}

// Final moments: sum the z-slices and multiply by i^|alpha|/alpha!.
__global__ void k_taylor_finish(int L, int C, int nz, const cpx<double>* part, const cpx<double>* calpha,
                                cpx<double>* mom) {
  const long long tot = (long long)L * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i / L), a = (int)(i % L);
    cpx<double> s = mk(0.0, 0.0);
    for (int zz = 0; zz < nz; ++zz) s = s + part[((long long)zz * C + c) * L + a];
    mom[(long long)c * L + a] = calpha[a] * s;  // column-major L x C
  }
}

// Evaluation: one thread per point, nested Horner over (a1, a2, a3).
__global__ void k_taylor_eval(long long M, const double* px, const double* py, const double* pz,
                              const int* cid, int p, int L, int C, const double* centers,
                              const cpx<double>* mom, cpx<double>* u) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < M;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = cid[i];
    if (c < 0 || c >= C) {  // invalid center id: flag the point instead of reading out of bounds
      u[i] = mk<double>(NAN, NAN);
      continue;
    }
    const double dx = px[i] - centers[3 * c], dy = py[i] - centers[3 * c + 1], dz = pz[i] - centers[3 * c + 2];
    const cpx<double>* mc = mom + (long long)c * L;
    // alpha index of (a1, a2, a3): terms are ordered a1 outer, a2, a3 inner
    cpx<double> acc1 = mk(0.0, 0.0);
    double x1 = 1.0;
    int idx = 0;
    for (int a1 = 0; a1 <= p; ++a1) {
      cpx<double> acc2 = mk(0.0, 0.0);
      double y2 = 1.0;
      for (int a2 = 0; a2 <= p - a1; ++a2) {
        cpx<double> acc3 = mk(0.0, 0.0);
        double z3 = 1.0;
        for (int a3 = 0; a3 <= p - a1 - a2; ++a3) {
          acc3 = acc3 + z3 * mc[idx++];
          z3 *= dz;
        }
        acc2 = acc2 + y2 * acc3;
        y2 *= dy;
      }
      acc1 = acc1 + x1 * acc2;
      x1 *= dx;
    }
    u[i] = acc1;
  }
}

// Direct sum u(r) = sum_j c_j exp(i q_j . r) (brute force, one thread per point).
__global__ void k_direct(long long NF, const double* qx, const double* qy, const double* qz,
                         const cpx<double>* coeff, long long M, const double* px, const double* py,
                         const double* pz, cpx<double>* u) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < M;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = px[i], y = py[i], z = pz[i];
    double sr = 0, si = 0;
    for (long long j = 0; j < NF; ++j) {
      double sn, cs;
      sincos(qx[j] * x + qy[j] * y + qz[j] * z, &sn, &cs);
      const cpx<double> cj = coeff[j];
      sr += cs * cj.re - sn * cj.im;
      si += cs * cj.im + sn * cj.re;
    }
    u[i] = mk(sr, si);
  }
}

static void alpha_tables(int p, std::vector<int>& al, std::vector<cpx<double>>& ca) {
  al.clear();
  ca.clear();
  for (int a1 = 0; a1 <= p; ++a1)
    for (int a2 = 0; a2 <= p - a1; ++a2)
      for (int a3 = 0; a3 <= p - a1 - a2; ++a3) {
        al.push_back(a1);
        al.push_back(a2);
        al.push_back(a3);
        const int s = a1 + a2 + a3;
        const double f = std::exp(-(std::lgamma(a1 + 1.0) + std::lgamma(a2 + 1.0) + std::lgamma(a3 + 1.0)));
        // i^s
        const int r = s % 4;
        const double re = (r == 0) ? 1.0 : (r == 2 ? -1.0 : 0.0);
        const double im = (r == 1) ? 1.0 : (r == 3 ? -1.0 : 0.0);
        ca.push_back(mk(re * f, im * f));
      }
}

}  // namespace fame6d

using namespace fame6d;

extern "C" {

int fame6d_taylor_moments(long long NF, const double* qmodes, const void* coeff, int p, int C,
                          const double* centers_host, void* moments, void* stream) {
  try {
    if (!qmodes || !coeff || !centers_host || !moments) throw Error(FAME6D_EARG, "null argument");
    if (p < 0 || p > kTaylorMaxP || C <= 0 || NF <= 0) throw Error(FAME6D_EARG, "bad Taylor sizes");
    hipStream_t s = (hipStream_t)stream;
    std::vector<int> al;
    std::vector<cpx<double>> ca;
    alpha_tables(p, al, ca);
    const int L = (int)ca.size();
    const int nz = (int)std::min<long long>(64, std::max<long long>(1, NF / 4096));
    DevBuf tabs, part;
    tabs.ensure(sizeof(int) * al.size() + sizeof(cpx<double>) * ca.size() + sizeof(double) * 3 * C + 64);
    char* b = tabs.as<char>();
    int* al_d = (int*)b;
    cpx<double>* ca_d = (cpx<double>*)(b + ((sizeof(int) * al.size() + 15) / 16) * 16);
    double* cen_d = (double*)(ca_d + ca.size());
    F6_HIP(hipMemcpyAsync(al_d, al.data(), sizeof(int) * al.size(), hipMemcpyHostToDevice, s));
    F6_HIP(hipMemcpyAsync(ca_d, ca.data(), sizeof(cpx<double>) * ca.size(), hipMemcpyHostToDevice, s));
    F6_HIP(hipMemcpyAsync(cen_d, centers_host, sizeof(double) * 3 * C, hipMemcpyHostToDevice, s));
    part.ensure(sizeof(cpx<double>) * (size_t)nz * C * L);
    dim3 grid(C, (L + kTaylorChunk - 1) / kTaylorChunk, nz);
    k_taylor_moments<<<grid, 256, 0, s>>>(NF, qmodes, qmodes + NF, qmodes + 2 * NF, (const cpx<double>*)coeff,
                                          p, cen_d, al_d, ca_d, L, part.as<cpx<double>>());
    F6_CHECK_LAUNCH();
    k_taylor_finish<<<grid_for((long long)L * C), 256, 0, s>>>(L, C, nz, part.as<cpx<double>>(), ca_d,
                                                               (cpx<double>*)moments);
    F6_CHECK_LAUNCH();
    F6_HIP(hipStreamSynchronize(s));
    tabs.release();
    part.release();
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

int fame6d_taylor_evaluate(long long M, const double* points, const int* center_id, int p, int C,
                           const double* centers_host, const void* moments, void* u, void* stream) {
  try {
    if (!points || !center_id || !centers_host || !moments || !u) throw Error(FAME6D_EARG, "null argument");
    if (p < 0 || p > kTaylorMaxP || C <= 0 || M < 0) throw Error(FAME6D_EARG, "bad Taylor sizes");
    hipStream_t s = (hipStream_t)stream;
    DevBuf cen;
    cen.ensure(sizeof(double) * 3 * C);
    F6_HIP(hipMemcpyAsync(cen.ptr, centers_host, sizeof(double) * 3 * C, hipMemcpyHostToDevice, s));
    const int L = taylor_terms(p);
    if (M > 0) {
      k_taylor_eval<<<grid_for(M), 256, 0, s>>>(M, points, points + M, points + 2 * M, center_id, p, L, C,
                                                 cen.as<double>(), (const cpx<double>*)moments,
                                                 (cpx<double>*)u);
      F6_CHECK_LAUNCH();
    }
    F6_HIP(hipStreamSynchronize(s));
    cen.release();
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

int fame6d_direct_reconstruct(long long NF, const double* qmodes, const void* coeff, long long M,
                              const double* points, void* u, void* stream) {
  try {
    if (!qmodes || !coeff || !points || !u) throw Error(FAME6D_EARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    if (M > 0) {
      k_direct<<<grid_for(M), 256, 0, s>>>(NF, qmodes, qmodes + NF, qmodes + 2 * NF, (const cpx<double>*)coeff,
                                            M, points, points + M, points + 2 * M, (cpx<double>*)u);
      F6_CHECK_LAUNCH();
    }
    F6_HIP(hipStreamSynchronize(s));
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return FAME6D_OK;
}

}  // extern "C"
