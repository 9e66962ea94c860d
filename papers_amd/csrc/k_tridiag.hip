// Eigenvalues of a Hermitian matrix by Householder tridiagonalisation and bisection (FP64).
//
// Used for the reduced Galerkin pencil C = W K W^H of the intrinsic Laplacian, where only the
// lowest eigenvalues are wanted. Step k (k = 0 .. n-2) takes the reflector H = I - tau v v^H
// with H^H (alpha, x) = (beta, 0), beta real, of the column below the diagonal (LAPACK's zlarfg
// convention), and applies H^H A H to the trailing block as the rank-2 update
//   A <- A - v w^H - w v^H,  p = tau A v,  w = p - (tau/2)(p^H v) v,
// leaving the real tridiagonal d (diagonal), e (beta_k). The Sturm count of the tridiagonal then
// brackets each wanted eigenvalue by bisection to the last bits. Two launches per step; the
// trailing matrix stays in the device caches between them.
#include "cy3q_internal.h"

namespace cy3q {

static constexpr int HH_THREADS = 256;
static constexpr int HH_ROWS = 4;   // rows of p per matvec workgroup (1 per wave)
static constexpr int HH_TILE = 64;  // update tile edge
static constexpr double kEps = 2.220446049250313e-16;  // 2^-52

__device__ inline double hh_block_sum(double v, double *red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int off = HH_THREADS / 2; off > 0; off >>= 1) {
    if (tid < off) red[tid] += red[tid + off];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ inline double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Reflector of step k and p = tau A_trail v. Every workgroup forms v (length m = n - k - 1) from
// column k below the diagonal, which the previous step's update (or k_hh_col0 for k = 0) left
// contiguous in col, in LDS (dynamic, m double2); workgroup 0 also publishes v, (beta, tau) and
// d[k], e[k]. With tau = 0 (x = 0 and alpha real) H = I, p = 0 and the update kernel does nothing.
__global__ void __launch_bounds__(HH_THREADS) k_hh_matvec(const double2 *__restrict__ A, int n,
                                                          int k, const double2 *__restrict__ col,
                                                          double2 *__restrict__ v_out,
                                                          double2 *__restrict__ p,
                                                          double *__restrict__ scal,
                                                          double *__restrict__ d,
                                                          double *__restrict__ e) {
  extern __shared__ double2 vs[];
  __shared__ double red[HH_THREADS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = n - k - 1;
  double xn = 0.0;
  for (int i = tid; i < m; i += HH_THREADS) {
    const double2 a = col[i];
    vs[i] = a;
    if (i > 0) xn += a.x * a.x + a.y * a.y;
  }
  const double xnorm2 = hh_block_sum(xn, red);  // also orders the LDS writes before the reads
  const double2 alpha = vs[0];
  double beta = alpha.x, tr = 0.0, ti = 0.0;
  double2 sc = make_double2(0.0, 0.0);
  if (xnorm2 > 0.0 || alpha.y != 0.0) {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tr = (beta - alpha.x) / beta;
    ti = -alpha.y / beta;
    // sc = 1 / (alpha - beta)
    const double ar = alpha.x - beta, ai = alpha.y, den = ar * ar + ai * ai;
    sc = make_double2(ar / den, -ai / den);
  }
  const bool ident = tr == 0.0 && ti == 0.0;
  __syncthreads();
  for (int i = tid; i < m; i += HH_THREADS)
    vs[i] = ident ? make_double2(i == 0 ? 1.0 : 0.0, 0.0)
                  : (i == 0 ? make_double2(1.0, 0.0) : cmul2(vs[i], sc));
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int i = tid; i < m; i += HH_THREADS) v_out[i] = vs[i];
    if (tid == 0) {
      scal[0] = beta;
      scal[1] = tr;
      scal[2] = ti;
      d[k] = A[(int64_t)k * n + k].x;
      e[k] = beta;
    }
  }
  // p_r = tau sum_j A[k+1+r][k+1+j] v_j, one wave per row.
  const int r0 = blockIdx.x * HH_ROWS;
  for (int rr = wave; rr < HH_ROWS; rr += HH_THREADS / 64) {
    const int r = r0 + rr;
    if (r >= m) break;
    const double2 *row = A + (int64_t)(k + 1 + r) * n + (k + 1);
    double sr = 0.0, si = 0.0;
    for (int j = lane; j < m; j += 64) {
      const double2 a = row[j], b = vs[j];
      sr += a.x * b.x - a.y * b.y;
      si += a.x * b.y + a.y * b.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_xor(sr, off);
      si += __shfl_xor(si, off);
    }
    if (lane == 0) p[r] = make_double2(tr * sr - ti * si, tr * si + ti * sr);
  }
}

// A_trail -= v w^H + w v^H with w = p + alpha v, alpha = -(tau/2)(p^H v); one 64 x 64 tile of the
// trailing block per workgroup (every workgroup forms alpha itself). The workgroups of tile
// column 0 also copy the updated column k + 1 below the diagonal into col (contiguous) for the
// next step's matvec.
__global__ void __launch_bounds__(HH_THREADS) k_hh_update(double2 *__restrict__ A, int n, int k,
                                                          const double2 *__restrict__ v,
                                                          const double2 *__restrict__ p,
                                                          const double *__restrict__ scal,
                                                          double2 *__restrict__ col) {
  __shared__ double red[HH_THREADS];
  __shared__ double2 vi[HH_TILE], wi[HH_TILE], vj[HH_TILE], wj[HH_TILE];
  const double tr = scal[1], ti = scal[2];
  const int tid = threadIdx.x;
  const int m = n - k - 1;
  if (tr == 0.0 && ti == 0.0) {  // H = I: the trailing block is unchanged
    if (blockIdx.x == 0)
      for (int q = tid; q < HH_TILE; q += HH_THREADS) {
        const int i = blockIdx.y * HH_TILE + q;
        if (i >= 1 && i < m) col[i - 1] = A[(int64_t)(k + 1 + i) * n + (k + 1)];
      }
    return;
  }
  double dr = 0.0, di = 0.0;  // p^H v
  for (int i = tid; i < m; i += HH_THREADS) {
    const double2 a = p[i], b = v[i];
    dr += a.x * b.x + a.y * b.y;
    di += a.x * b.y - a.y * b.x;
  }
  dr = hh_block_sum(dr, red);
  di = hh_block_sum(di, red);
  // alpha = -0.5 tau (p^H v)
  const double alr = -0.5 * (tr * dr - ti * di), ali = -0.5 * (tr * di + ti * dr);
  const int i0 = blockIdx.y * HH_TILE, j0 = blockIdx.x * HH_TILE;
  if (tid < HH_TILE) {
    const int i = i0 + tid;
    if (i < m) {
      const double2 vv = v[i], pp = p[i];
      vi[tid] = vv;
      wi[tid] = make_double2(pp.x + alr * vv.x - ali * vv.y, pp.y + alr * vv.y + ali * vv.x);
    }
  } else if (tid < 2 * HH_TILE) {
    const int j = j0 + tid - HH_TILE;
    if (j < m) {
      const double2 vv = v[j], pp = p[j];
      vj[tid - HH_TILE] = vv;
      wj[tid - HH_TILE] = make_double2(pp.x + alr * vv.x - ali * vv.y, pp.y + alr * vv.y + ali * vv.x);
    }
  }
  __syncthreads();
  for (int q = tid; q < HH_TILE * HH_TILE; q += HH_THREADS) {
    const int ii = q / HH_TILE, jj = q % HH_TILE;
    const int i = i0 + ii, j = j0 + jj;
    if (i >= m || j >= m) continue;
    // v_i conj(w_j) + w_i conj(v_j)
    const double2 a = vi[ii], b = wj[jj], c = wi[ii], f = vj[jj];
    const double ur = a.x * b.x + a.y * b.y + c.x * f.x + c.y * f.y;
    const double ui = a.y * b.x - a.x * b.y + c.y * f.x - c.x * f.y;
    double2 *x = A + (int64_t)(k + 1 + i) * n + (k + 1 + j);
    const double2 old = *x;
    const double2 nv = make_double2(old.x - ur, old.y - ui);
    *x = nv;
    if (j == 0 && i >= 1) col[i - 1] = nv;
  }
}

// Column 0 below the diagonal, contiguous, for step 0.
__global__ void k_hh_col0(const double2 *__restrict__ A, int n, double2 *__restrict__ col) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n - 1) col[i] = A[(int64_t)(i + 1) * n];
}

__global__ void k_hh_last(const double2 *__restrict__ A, int n, double *__restrict__ d) {
  d[n - 1] = A[(int64_t)(n - 1) * n + (n - 1)].x;
}

// The j-th smallest eigenvalue (j < n_want) of the symmetric tridiagonal (d, e) by multisection on
// the Sturm count, from the Gershgorin interval down to a few ulps: one wave per eigenvalue, whose
// 64 lanes count at 64 interior points of the current bracket, so each round narrows it 65-fold
// (about 10 rounds instead of about 60 halvings, each a serial pass over the n pivots).
__device__ inline int sturm_count(const double *__restrict__ d, const double *__restrict__ e, int n,
                                  double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (q < 0.0) cnt++;
  for (int i = 1; i < n; i++) {
    if (fabs(q) < pivmin) q = -pivmin;
    q = (d[i] - x) - e[i - 1] * e[i - 1] / q;
    if (q < 0.0) cnt++;
  }
  return cnt;
}

__global__ void __launch_bounds__(64) k_tridiag_bisect(const double *__restrict__ d,
                                                       const double *__restrict__ e, int n,
                                                       int n_want, double *__restrict__ ev) {
  const int j = blockIdx.x, lane = threadIdx.x;
  if (j >= n_want) return;
  double lo = INFINITY, hi = -INFINITY, emax2 = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    if (i < n - 1) emax2 = fmax(emax2, e[i] * e[i]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off));
    hi = fmax(hi, __shfl_xor(hi, off));
    emax2 = fmax(emax2, __shfl_xor(emax2, off));
  }
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax2);
  const double span = fmax(fabs(lo), fabs(hi));
  lo -= 2.0 * kEps * span + pivmin;
  hi += 2.0 * kEps * span + pivmin;
  // Invariant: count(lo) <= j < count(hi). Every lane holds the same bracket.
  for (int it = 0; it < 64; it++) {
    if (hi - lo <= 2.0 * kEps * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    const double x = lo + (hi - lo) * ((double)(lane + 1) / 65.0);
    const bool above = sturm_count(d, e, n, x, pivmin) > j;
    const unsigned long long mask = __ballot(above);
    const int first = mask ? __ffsll((long long)mask) - 1 : 64;  // first point with count > j
    const double xlo = __shfl(x, first > 0 ? first - 1 : 0);
    const double xhi = __shfl(x, first < 64 ? first : 63);
    const double nlo = first > 0 ? xlo : lo, nhi = first < 64 ? xhi : hi;
    if (nlo == lo && nhi == hi) break;  // the points no longer split the bracket
    lo = nlo;
    hi = nhi;
  }
  if (lane == 0) ev[j] = 0.5 * (lo + hi);
}

int tridiag_smem_ok(int n) { return (size_t)n * sizeof(double2) <= 160 * 1024 - 4096; }

// v and p hold n complex values each; p[n .. 2n) is the contiguous column buffer.
void launch_hh_tridiag(hipStream_t st, double2 *A, int n, double2 *v, double2 *p, double *scal,
                       double *d, double *e) {
  double2 *col = p + n;
  if (n > 1)
    hipLaunchKernelGGL(k_hh_col0, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A, n, col);
  for (int k = 0; k + 1 < n; k++) {
    const int m = n - k - 1;
    hipLaunchKernelGGL(k_hh_matvec, dim3((unsigned)((m + HH_ROWS - 1) / HH_ROWS)),
                       dim3(HH_THREADS), (size_t)m * sizeof(double2), st, A, n, k, col, v, p,
                       scal, d, e);
    const unsigned nt = (unsigned)((m + HH_TILE - 1) / HH_TILE);
    hipLaunchKernelGGL(k_hh_update, dim3(nt, nt), dim3(HH_THREADS), 0, st, A, n, k, v, p, scal,
                       col);
  }
  hipLaunchKernelGGL(k_hh_last, dim3(1), dim3(1), 0, st, A, n, d);
}

void launch_tridiag_bisect(hipStream_t st, const double *d, const double *e, int n, int n_want,
                           double *ev) {
  if (n_want <= 0) return;
  hipLaunchKernelGGL(k_tridiag_bisect, dim3((unsigned)n_want), dim3(64), 0, st, d, e,
                     n, n_want, ev);
}

}  // namespace cy3q
