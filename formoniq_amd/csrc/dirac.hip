// Hodge-Dirac operator on the device and its source problem.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   HodgeDirac::assemble_signed   formoniq/src/problems/dirac.rs:213-258
//   solve_dirac_source            formoniq/src/problems/dirac.rs:473-534
//   top_harmonic                  formoniq/src/problems/dirac.rs:553-568
//   solve_bordered                formoniq/src/problems/dirac.rs:578-598
//   Leapfrog::{new, step, conserved_energy}   formoniq/src/time.rs:292-420
//   LinearIrk::{new, step}, stage_matrix      formoniq/src/time.rs:191-274 (Gauss-Legendre)
//   solve_dirac_leapfrog, grade_parity_coloring  formoniq/src/problems/dirac.rs:330-337, 415-434
//
// mass_block is the block diagonal of the grade masses M_k; op carries
// U_k = M_k D_{k-1} (global CSR x CSR product, as the reference) below the
// diagonal and sign * U_k^T above it. The source problem (the reference
// factors it with sparse LU) is solved here with MINRES on the symmetric
// (self-adjoint sign) system, preconditioned by the inverse absolute diagonal
// of the mass block, bordered with the top-grade harmonic when m = 0.
#include <algorithm>
#include <cmath>
#include <vector>

#include "feec_internal.hpp"

namespace feec {

void assemble_form(Geom& g, int bc, int kind, int grade, Csr& out);

namespace {
constexpr int TPB = 256;
inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }

__global__ void k_abs_inv(const double* d, int64_t n, double* out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = d[i] != 0.0 ? 1.0 / fabs(d[i]) : 1.0;
}

// Flat field offsets of the grades 0..ngrades-1 (ngrades + 1 entries).
struct GradeOffsets {
  int64_t off[6];
  int ngrades;
};

// out = a * b (elementwise)
__global__ void k_mul(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                      double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] * b[i];
}

// x = s a
__global__ void k_scaled(const double* __restrict__ a, int64_t n, double s, double* __restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = s * a[i];
}

// Sum of a block's values on thread 0 (fixed order: wavefront tree, then the
// block's wavefronts in order).
__device__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (unsigned w = 0; w < blockDim.x / 64; ++w) t += sh[w];
  return t;
}

// part[i * gridDim.x + b] = block b's share of V_i . w (blockIdx.y = i)
__global__ void __launch_bounds__(TPB) k_multi_dot(const double* __restrict__ V, int64_t stride,
                                                   const double* __restrict__ w, int64_t n,
                                                   double* __restrict__ part) {
  __shared__ double sh[TPB / 64];
  const double* v = V + blockIdx.y * stride;
  double acc = 0.0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x)
    acc += v[r] * w[r];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = t;
}

// h[i] (+)= sum of the np partials of i (block i); accumulate when `add`
__global__ void __launch_bounds__(TPB) k_multi_finish(const double* __restrict__ part, int np,
                                                      double* __restrict__ h, int add) {
  __shared__ double sh[TPB / 64];
  double acc = 0.0;
  for (int b = threadIdx.x; b < np; b += blockDim.x) acc += part[blockIdx.x * np + b];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) h[blockIdx.x] = add ? h[blockIdx.x] + t : t;
}

// w -= sum_i c[i] V_i, with c on the device (nv terms)
__global__ void k_multi_axpy(const double* __restrict__ V, int64_t stride, int nv,
                             const double* __restrict__ c, int64_t n, double* __restrict__ w) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  double acc = w[r];
  for (int i = 0; i < nv; ++i) acc -= c[i] * V[i * stride + r];
  w[r] = acc;
}

// x = y on the rows whose grade has parity `odd`, 0 on the others (the
// grade-parity coloring of the leapfrog split).
__global__ void k_parity_part(const double* __restrict__ y, GradeOffsets go, int odd,
                              double* __restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= go.off[go.ngrades]) return;
  int k = 0;
  while (k + 1 < go.ngrades && i >= go.off[k + 1]) ++k;
  x[i] = (k & 1) == odd ? y[i] : 0.0;
}
}  // namespace

struct Dirac {
  std::vector<int64_t> offsets;
  std::vector<Csr> masses;
  Csr mass_block, op;
};

static void coboundary_bc(Mesh& m, int bc, int k, Csr& out) {
  out.ctx = m.ctx;
  if (bc == FEEC_RELATIVE) {
    Csr full;
    mesh_coboundary(m, k, full);
    m.ensure_interior();
    csr_select(full, m.interior[k + 1].p, m.n_interior[k + 1], m.interior[k].p, m.n_interior[k],
               out);
  } else {
    mesh_coboundary(m, k, out);
  }
}

void dirac_assemble(Geom& g, int bc, double sign, Dirac& D) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  const int d = m.d;
  D.masses.resize(d + 1);
  D.offsets.assign(d + 2, 0);
  for (int k = 0; k <= d; ++k) {
    assemble_form(g, bc, FEEC_MASS, k, D.masses[k]);
    D.offsets[k + 1] = D.offsets[k] + D.masses[k].nrows;
  }
  std::vector<Csr> U(d + 1);
  for (int k = 1; k <= d; ++k) {
    Csr dif;
    coboundary_bc(m, bc, k - 1, dif);
    csr_mul(D.masses[k], dif, U[k]);
  }
  const int nb = d + 1;
  std::vector<int64_t> sizes(nb);
  for (int k = 0; k <= d; ++k) sizes[k] = D.masses[k].nrows;
  std::vector<BlockRef> mb(nb * nb, BlockRef{nullptr, 1.0, false});
  for (int k = 0; k <= d; ++k) mb[k * nb + k] = BlockRef{&D.masses[k], 1.0, false};
  csr_block(ctx, nb, mb.data(), sizes.data(), sizes.data(), D.mass_block);
  std::vector<BlockRef> ob(nb * nb, BlockRef{nullptr, 1.0, false});
  for (int k = 1; k <= d; ++k) {
    ob[k * nb + (k - 1)] = BlockRef{&U[k], 1.0, false};
    ob[(k - 1) * nb + k] = BlockRef{&U[k], sign, true};
  }
  csr_block(ctx, nb, ob.data(), sizes.data(), sizes.data(), D.op);
}

}  // namespace feec

using namespace feec;

extern "C" {

feec_status feec_hodge_dirac_assemble(feec_geom* g, feec_bc bc, double sign, feec_csr** mass_block,
                                      feec_csr** op, int64_t* offsets) {
  return guard([&] {
    FEEC_REQUIRE(g, FEEC_E_ARG, "null geometry");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    FEEC_HIP(hipSetDevice(g->mesh->ctx->device));
    Dirac D;
    dirac_assemble(*g, bc, sign, D);
    if (offsets)
      for (size_t i = 0; i < D.offsets.size(); ++i) offsets[i] = D.offsets[i];
    if (mass_block) {
      auto* a = new feec_csr_s();
      static_cast<Csr&>(*a) = std::move(D.mass_block);
      *mass_block = a;
    }
    if (op) {
      auto* a = new feec_csr_s();
      static_cast<Csr&>(*a) = std::move(D.op);
      *op = a;
    }
  });
}

// Explicit leapfrog of the Hodge-Dirac flow M y' = A y (A the skew op):
// grades of even parity are the drifted block q, odd ones the kicked block p;
// one step is p += dt/2 M^-1 A q, q += dt M^-1 A p, p += dt/2 M^-1 A q. The
// reference factors the two color-block masses; here each mass solve is a
// Jacobi CG (scalars on the device) on the whole block-diagonal mass matrix
// with a right-hand side supported on one color (the iterates stay on that
// color exactly). The
// closing half-kick's M^-1 A q is the next step's opening one, so a step
// takes two solves. energy[s] is the staggered invariant
// 1/2 y^T M y - dt^2/8 (A q)^T M^-1 (A q) of the state after s steps.
static void dirac_leapfrog(Geom& g, int bc, double dt, int64_t nsteps, double* y_host,
                           double* states, double* energy, feec_report* report, double cg_rtol) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  hipStream_t st = ctx->stream;
  struct Events {  // destroyed on every path out, errors included
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } ev;
  FEEC_HIP(hipEventCreate(&ev.e0));
  FEEC_HIP(hipEventCreate(&ev.e1));
  FEEC_HIP(hipEventRecord(ev.e0, st));
  Dirac D;
  dirac_assemble(g, bc, -1.0, D);
  const int d = m.d;
  const int64_t N = D.offsets[d + 1];
  GradeOffsets go{};
  go.ngrades = d + 1;
  for (int k = 0; k <= d + 1; ++k) go.off[k] = D.offsets[k];
  DBuf<double> y(ctx, N), x(ctx, N), r(ctx, N), z(ctx, N), my(ctx, N), dinv(ctx, N),
      part(ctx, 2 * 4096);
  copy_to_device(ctx, y.p, y_host, N * sizeof(double));
  csr_diagonal(D.mass_block, dinv.p);
  k_abs_inv<<<grid_for(N), TPB, 0, st>>>(dinv.p, N, dinv.p);
  FEEC_HIP(hipGetLastError());
  int64_t cg_iters = 0;
  double worst = 0.0;
  // r = A (parity part of y), z = M^-1 r
  auto kick = [&](int odd) {
    k_parity_part<<<grid_for(N), TPB, 0, st>>>(y.p, go, odd, x.p);
    spmv64(ctx, D.op, x.p, r.p);
    KrylovStats ks = cg_solve_device(ctx, D.mass_block, dinv.p, r.p, z.p, cg_rtol, 100000);
    FEEC_REQUIRE(ks.converged, FEEC_E_NOT_CONVERGED, "leapfrog mass solve did not converge");
    cg_iters += ks.iters;
    worst = std::max(worst, ks.precond_residual);
  };
  auto staggered_energy = [&]() {  // needs r = A q, z = M^-1 r of the current y
    spmv64(ctx, D.mass_block, y.p, my.p);
    return 0.5 * dot64(ctx, y.p, my.p, N, part.p) - dt * dt / 8.0 * dot64(ctx, r.p, z.p, N, part.p);
  };
  kick(0);
  if (energy) energy[0] = staggered_energy();
  if (states) std::copy(y_host, y_host + N, states);
  for (int64_t s = 0; s < nsteps; ++s) {
    axpy64(ctx, z.p, N, 0.5 * dt, y.p);  // p half-kick (z lives on the odd grades)
    kick(1);
    axpy64(ctx, z.p, N, dt, y.p);        // q drift (z on the even grades)
    kick(0);
    axpy64(ctx, z.p, N, 0.5 * dt, y.p);  // p half-kick
    if (energy) energy[s + 1] = staggered_energy();
    if (states) copy_to_host(ctx, states + (s + 1) * N, y.p, N * sizeof(double));
  }
  copy_to_host(ctx, y_host, y.p, N * sizeof(double));
  FEEC_HIP(hipEventRecord(ev.e1, st));
  FEEC_HIP(hipEventSynchronize(ev.e1));
  float ms = 0.f;
  FEEC_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
  if (report) {
    *report = feec_report{};
    report->iters = cg_iters;
    report->converged = 1;
    report->precond_residual = worst;
    report->seconds = ms * 1e-3;
  }
}

// Gauss-Legendre collocation tableau with s = 1 (implicit midpoint) or s = 2
// stages: nodes 1/2 -+ sqrt(3)/6, A = [[1/4, 1/4 - sqrt(3)/6], [1/4 + sqrt(3)/6, 1/4]],
// weights 1/2 (the textbook coefficients of the 4th-order Gauss method).
struct GaussTableau {
  int s;
  double a[2][2], b[2];
};
static GaussTableau gauss_legendre(int s) {
  GaussTableau t{};
  t.s = s;
  if (s == 1) {
    t.a[0][0] = 0.5;
    t.b[0] = 1.0;
  } else {
    const double r = std::sqrt(3.0) / 6.0;
    t.a[0][0] = 0.25;
    t.a[0][1] = 0.25 - r;
    t.a[1][0] = 0.25 + r;
    t.a[1][1] = 0.25;
    t.b[0] = t.b[1] = 0.5;
  }
  return t;
}

// Implicit Gauss-Legendre steps of M y' = A y (LinearIrk::step with zero
// forcing): the stage derivatives K solve (I_s (x) M - dt (A_tab (x) A)) K =
// 1_s (x) A y0, then y1 = y0 + dt sum_i b_i K_i. The reference factors the
// stage matrix once; here each step solves it by restarted GMRES (right
// preconditioned by the inverse diagonal of M on every stage block; classical
// Gram-Schmidt applied twice, each pass one launch of batched dot products and
// one update), with the Hessenberg least-squares problem on the host and every
// vector on the device.
// energy[s] = 1/2 y^T M y of the state after s steps (conserved by the method).
static void dirac_irk(Geom& g, int bc, int stages, double dt, int64_t nsteps, double* y_host,
                      double* states, double* energy, feec_report* report, double rtol) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  hipStream_t st = ctx->stream;
  struct Events {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } ev;
  FEEC_HIP(hipEventCreate(&ev.e0));
  FEEC_HIP(hipEventCreate(&ev.e1));
  FEEC_HIP(hipEventRecord(ev.e0, st));
  Dirac D;
  dirac_assemble(g, bc, -1.0, D);
  const int64_t N = D.offsets[m.d + 1];
  const GaussTableau tab = gauss_legendre(stages);
  const int S = tab.s;
  const int64_t SN = S * N;
  constexpr int RESTART = 40;
  DBuf<double> y(ctx, N), my(ctx, N), ay(ctx, N), dinv(ctx, N), rhs(ctx, SN), k(ctx, SN),
      w(ctx, SN), t(ctx, SN), u(ctx, SN), part(ctx, 2 * 4096);
  DBuf<double> V(ctx, (RESTART + 1) * SN);
  const unsigned GD = std::min<unsigned>(grid_for(SN), 512);
  DBuf<double> mpart(ctx, (size_t)(RESTART + 1) * GD), hcol(ctx, 2 * (RESTART + 1));
  copy_to_device(ctx, y.p, y_host, N * sizeof(double));
  csr_diagonal(D.mass_block, dinv.p);
  k_abs_inv<<<grid_for(N), TPB, 0, st>>>(dinv.p, N, dinv.p);
  FEEC_HIP(hipGetLastError());
  // out = (I (x) M - dt A_tab (x) A) in, on S stacked blocks of N
  auto stage_apply = [&](const double* in, double* out) {
    for (int j = 0; j < S; ++j) spmv64(ctx, D.op, in + j * N, t.p + j * N);
    for (int i = 0; i < S; ++i) {
      spmv64(ctx, D.mass_block, in + i * N, out + i * N);
      for (int j = 0; j < S; ++j)
        if (tab.a[i][j] != 0.0) axpy64(ctx, t.p + j * N, N, -dt * tab.a[i][j], out + i * N);
    }
  };
  // u = P^-1 v: the inverse mass diagonal on every stage block
  auto precond = [&](const double* v, double* out) {
    for (int i = 0; i < S; ++i)
      k_mul<<<grid_for(N), TPB, 0, st>>>(dinv.p, v + i * N, N, out + i * N);
  };
  auto dot = [&](const double* a, const double* b) { return dot64(ctx, a, b, SN, part.p); };
  int64_t total_iters = 0;
  double worst = 0.0;
  auto gmres = [&]() {  // k = stage^-1 rhs from k = 0
    FEEC_HIP(hipMemsetAsync(k.p, 0, SN * sizeof(double), st));
    const double bnorm = std::sqrt(dot(rhs.p, rhs.p));
    if (bnorm == 0.0) return;
    double res = 1.0;
    for (int cycle = 0; cycle < 200 && res > rtol; ++cycle) {
      // r = rhs - S k into V_0
      stage_apply(k.p, w.p);
      FEEC_HIP(hipMemcpyAsync(V.p, rhs.p, SN * sizeof(double), hipMemcpyDeviceToDevice, st));
      axpy64(ctx, w.p, SN, -1.0, V.p);
      const double beta = std::sqrt(dot(V.p, V.p));
      res = beta / bnorm;
      if (res <= rtol) break;
      k_scaled<<<grid_for(SN), TPB, 0, st>>>(V.p, SN, 1.0 / beta, V.p);
      std::vector<double> H((RESTART + 1) * RESTART, 0.0), cs(RESTART), sn(RESTART),
          gv(RESTART + 1, 0.0);
      gv[0] = beta;
      int jn = 0;
      for (int j = 0; j < RESTART; ++j) {
        double* vj = V.p + j * SN;
        double* vn = V.p + (j + 1) * SN;
        precond(vj, u.p);
        stage_apply(u.p, vn);
        // classical Gram-Schmidt, twice (one dot launch and one update launch
        // per pass over V_0..V_j; the coefficients stay on the device)
        const int nv = j + 1;
        for (int pass = 0; pass < 2; ++pass) {
          k_multi_dot<<<dim3(GD, nv), TPB, 0, st>>>(V.p, SN, vn, SN, mpart.p);
          k_multi_finish<<<nv, TPB, 0, st>>>(mpart.p, (int)GD, hcol.p + pass * (RESTART + 1), 0);
          k_multi_axpy<<<grid_for(SN), TPB, 0, st>>>(V.p, SN, nv, hcol.p + pass * (RESTART + 1), SN,
                                                     vn);
        }
        FEEC_HIP(hipGetLastError());
        std::vector<double> hh(2 * (RESTART + 1));
        copy_to_host(ctx, hh.data(), hcol.p, hh.size() * sizeof(double));
        for (int i = 0; i <= j; ++i) H[i * RESTART + j] = hh[i] + hh[RESTART + 1 + i];
        const double hn = std::sqrt(dot(vn, vn));
        H[(j + 1) * RESTART + j] = hn;
        if (hn > 0.0) k_scaled<<<grid_for(SN), TPB, 0, st>>>(vn, SN, 1.0 / hn, vn);
        for (int i = 0; i < j; ++i) {  // earlier rotations on the new column
          const double a0 = H[i * RESTART + j], a1 = H[(i + 1) * RESTART + j];
          H[i * RESTART + j] = cs[i] * a0 + sn[i] * a1;
          H[(i + 1) * RESTART + j] = -sn[i] * a0 + cs[i] * a1;
        }
        const double a0 = H[j * RESTART + j], a1 = H[(j + 1) * RESTART + j];
        const double rr = std::hypot(a0, a1);
        cs[j] = rr > 0.0 ? a0 / rr : 1.0;
        sn[j] = rr > 0.0 ? a1 / rr : 0.0;
        H[j * RESTART + j] = rr;
        H[(j + 1) * RESTART + j] = 0.0;
        gv[j + 1] = -sn[j] * gv[j];
        gv[j] = cs[j] * gv[j];
        jn = j + 1;
        ++total_iters;
        res = std::fabs(gv[j + 1]) / bnorm;
        if (res <= rtol || hn == 0.0) break;
      }
      // z = H^-1 g (upper triangular), k += P^-1 sum_i z_i V_i
      std::vector<double> z(jn, 0.0);
      for (int i = jn - 1; i >= 0; --i) {
        double v = gv[i];
        for (int c = i + 1; c < jn; ++c) v -= H[i * RESTART + c] * z[c];
        z[i] = v / H[i * RESTART + i];
      }
      FEEC_HIP(hipMemsetAsync(w.p, 0, SN * sizeof(double), st));
      for (int i = 0; i < jn; ++i) axpy64(ctx, V.p + i * SN, SN, z[i], w.p);
      precond(w.p, u.p);
      axpy64(ctx, u.p, SN, 1.0, k.p);
    }
    FEEC_REQUIRE(res <= rtol, FEEC_E_NOT_CONVERGED, "Gauss-Legendre stage solve did not converge");
    worst = std::max(worst, res);
  };
  auto energy_of = [&]() {
    spmv64(ctx, D.mass_block, y.p, my.p);
    return 0.5 * dot64(ctx, y.p, my.p, N, part.p);
  };
  if (energy) energy[0] = energy_of();
  if (states) std::copy(y_host, y_host + N, states);
  for (int64_t step = 0; step < nsteps; ++step) {
    spmv64(ctx, D.op, y.p, ay.p);
    for (int i = 0; i < S; ++i)
      FEEC_HIP(hipMemcpyAsync(rhs.p + i * N, ay.p, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    gmres();
    for (int i = 0; i < S; ++i) axpy64(ctx, k.p + i * N, N, dt * tab.b[i], y.p);
    if (energy) energy[step + 1] = energy_of();
    if (states) copy_to_host(ctx, states + (step + 1) * N, y.p, N * sizeof(double));
  }
  copy_to_host(ctx, y_host, y.p, N * sizeof(double));
  FEEC_HIP(hipEventRecord(ev.e1, st));
  FEEC_HIP(hipEventSynchronize(ev.e1));
  float ms = 0.f;
  FEEC_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
  if (report) {
    *report = feec_report{};
    report->iters = total_iters;
    report->converged = 1;
    report->residual = worst;
    report->seconds = ms * 1e-3;
  }
}

feec_status feec_dirac_irk(feec_geom* g, feec_bc bc, int stages, double dt, int64_t nsteps,
                           double* y, double* states, double* energy, feec_report* report) {
  return guard([&] {
    FEEC_REQUIRE(g && y, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    FEEC_REQUIRE(stages == 1 || stages == 2, FEEC_E_ARG, "stages must be 1 or 2");
    FEEC_REQUIRE(nsteps >= 0 && std::isfinite(dt), FEEC_E_ARG, "nsteps >= 0 and a finite dt");
    FEEC_HIP(hipSetDevice(g->mesh->ctx->device));
    dirac_irk(*g, bc, stages, dt, nsteps, y, states, energy, report, 1e-14);
  });
}

feec_status feec_dirac_leapfrog(feec_geom* g, feec_bc bc, double dt, int64_t nsteps, double* y,
                                double* states, double* energy, feec_report* report) {
  return guard([&] {
    FEEC_REQUIRE(g && y, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    FEEC_REQUIRE(nsteps >= 0 && std::isfinite(dt), FEEC_E_ARG, "nsteps >= 0 and a finite dt");
    FEEC_HIP(hipSetDevice(g->mesh->ctx->device));
    dirac_leapfrog(*g, bc, dt, nsteps, y, states, energy, report, 1e-14);
  });
}

feec_status feec_solve_dirac_source(feec_geom* g, double mass_term, const double* load,
                                    const double* lift, const feec_solve_opts* opts, double* out,
                                    feec_report* report) {
  return guard([&] {
    FEEC_REQUIRE(g && load && out, FEEC_E_ARG, "null geometry, load or output");
    FEEC_REQUIRE(std::isfinite(mass_term), FEEC_E_ARG, "mass_term must be finite");
    Mesh& m = *g->mesh;
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    feec_solve_opts o;
    if (opts)
      o = *opts;
    else
      feec_solve_opts_default(&o);
    o.precision = FEEC_PREC_FP64;  // the reference solve is a direct FP64 factorization
    const int d = m.d;
    int64_t nflat = 0;
    for (int k = 0; k <= d; ++k) nflat += m.n[k];
    for (int64_t i = 0; i < nflat; ++i)
      FEEC_REQUIRE(std::isfinite(load[i]) && (!lift || std::isfinite(lift[i])), FEEC_E_ARG,
                   "load and lift must be finite");
    Dirac D;
    dirac_assemble(*g, FEEC_FULL, 1.0, D);
    const int64_t nfull = D.offsets[d + 1];
    // S = A + m M (full), as one COO sum
    Csr S;
    if (mass_term != 0.0) {
      int64_t sz[1] = {nfull};
      BlockRef scaled{&D.mass_block, mass_term, false};
      Csr mm;
      csr_block(ctx, 1, &scaled, sz, sz, mm);
      csr_add(D.op, mm, S);
    } else {
      int64_t sz[1] = {nfull};
      BlockRef same{&D.op, 1.0, false};
      csr_block(ctx, 1, &same, sz, sz, S);
    }
    // the flat inclusion E over all grades
    m.ensure_interior();
    std::vector<int32_t> map(nfull);
    int64_t nrel = 0;
    for (int k = 0; k <= d; ++k) {
      std::vector<int32_t> ik(m.n[k]);
      copy_to_host(ctx, ik.data(), m.interior[k].p, m.n[k] * 4);
      for (int64_t i = 0; i < m.n[k]; ++i)
        map[D.offsets[k] + i] = ik[i] >= 0 ? (int32_t)(nrel + ik[i]) : -1;
      nrel += m.n_interior[k];
    }
    DBuf<int32_t> dmap(ctx, nfull);
    copy_to_device(ctx, dmap.p, map.data(), nfull * 4);
    // rhs = E^T (f - S lift)
    std::vector<double> f(load, load + nfull), l(nfull, 0.0), sl(nfull);
    if (lift) std::copy(lift, lift + nfull, l.begin());
    DBuf<double> dl(ctx, nfull), dsl(ctx, nfull);
    copy_to_device(ctx, dl.p, l.data(), nfull * 8);
    spmv64(ctx, S, dl.p, dsl.p);
    ctx->sync();
    copy_to_host(ctx, sl.data(), dsl.p, nfull * 8);
    Csr Srel;
    csr_select(S, dmap.p, nrel, dmap.p, nrel, Srel);
    std::vector<double> rhs(nrel, 0.0);
    for (int64_t i = 0; i < nfull; ++i)
      if (map[i] >= 0) rhs[map[i]] = f[i] - sl[i];
    // m = 0: border with the top harmonic h = M_n^-1 o when it is in the kernel
    std::vector<double> h;
    bool bordered = false;
    if (mass_term == 0.0) {
      m.ensure_orientation();
      if (m.orientable) {
        Csr Mn;
        assemble_form(*g, FEEC_RELATIVE, FEEC_MASS, d, Mn);
        std::vector<double> dg(Mn.nrows);
        DBuf<double> dd(ctx, Mn.nrows);
        csr_diagonal(Mn, dd.p);
        ctx->sync();
        copy_to_host(ctx, dg.data(), dd.p, Mn.nrows * 8);
        h.assign(nrel, 0.0);
        int64_t off = nrel - m.n_interior[d];
        for (int64_t c = 0; c < m.n_interior[d]; ++c) h[off + c] = m.orientation[c] / dg[c];
        DBuf<double> dh(ctx, nrel), sh(ctx, nrel);
        copy_to_device(ctx, dh.p, h.data(), nrel * 8);
        spmv64(ctx, Srel, dh.p, sh.p);
        ctx->sync();
        std::vector<double> shh(nrel);
        copy_to_host(ctx, shh.data(), sh.p, nrel * 8);
        double a = 0.0, b = 0.0;
        for (int64_t i = 0; i < nrel; ++i) {
          a += shh[i] * shh[i];
          b += h[i] * h[i];
        }
        bordered = std::sqrt(a) <= 1e-8 * std::sqrt(b);
      }
    }
    KktSystem sys;
    sys.ctx = ctx;
    int64_t n = nrel + (bordered ? 1 : 0);
    sys.n = n;
    sys.nh = 0;
    if (bordered) {
      // [[S, h], [h^T, 0]] with the nonzero entries of h (solve_bordered)
      std::vector<int64_t> rp(nrel + 1), rp2(2);
      std::vector<int32_t> c1, c2;
      std::vector<double> v1, v2;
      for (int64_t i = 0; i < nrel; ++i) {
        rp[i] = (int64_t)c1.size();
        if (h[i] != 0.0) {
          c1.push_back(0);
          v1.push_back(h[i]);
          c2.push_back((int32_t)i);
          v2.push_back(h[i]);
        }
      }
      rp[nrel] = (int64_t)c1.size();
      rp2[0] = 0;
      rp2[1] = (int64_t)c2.size();
      auto up = [&](Csr& c, int64_t r, int64_t cc, std::vector<int64_t>& rpv,
                    std::vector<int32_t>& cv, std::vector<double>& vv) {
        c.ctx = ctx;
        c.nrows = r;
        c.ncols = cc;
        c.nnz = (int64_t)cv.size();
        c.row_ptr.alloc(ctx, r + 1);
        c.col.alloc(ctx, c.nnz);
        c.val.alloc(ctx, c.nnz);
        copy_to_device(ctx, c.row_ptr.p, rpv.data(), (r + 1) * 8);
        copy_to_device(ctx, c.col.p, cv.data(), c.nnz * 4);
        copy_to_device(ctx, c.val.p, vv.data(), c.nnz * 8);
      };
      Csr bc_col, bc_row;
      up(bc_col, nrel, 1, rp, c1, v1);
      up(bc_row, 1, nrel, rp2, c2, v2);
      BlockRef grid[4] = {{&Srel, 1.0, false}, {&bc_col, 1.0, false}, {&bc_row, 1.0, false},
                          {nullptr, 1.0, false}};
      int64_t rs[2] = {nrel, 1};
      csr_block(ctx, 2, grid, rs, rs, sys.A);
      rhs.push_back(0.0);
    } else {
      sys.A = std::move(Srel);
    }
    // preconditioner: inverse absolute diagonal of the relative mass block
    Csr Mrel;
    csr_select(D.mass_block, dmap.p, nrel, dmap.p, nrel, Mrel);
    DBuf<double> md(ctx, n);
    FEEC_HIP(hipMemsetAsync(md.p, 0, n * 8, ctx->stream));
    csr_diagonal(Mrel, md.p);
    sys.inv_diag.alloc(ctx, n);
    k_abs_inv<<<grid_for(n), TPB, 0, ctx->stream>>>(md.p, n, sys.inv_diag.p);
    FEEC_HIP(hipGetLastError());
    DBuf<double> b(ctx, n), x(ctx, n);
    ctx->sync();
    copy_to_device(ctx, b.p, rhs.data(), n * 8);
    SolverState* ws = nullptr;
    feec_report r = kkt_solve(sys, b.p, x.p, o, &ws);
    solver_state_free(ws);
    if (report) *report = r;
    std::vector<double> xs(n);
    copy_to_host(ctx, xs.data(), x.p, n * 8);
    for (int64_t i = 0; i < nfull; ++i) out[i] = l[i] + (map[i] >= 0 ? xs[map[i]] : 0.0);
    FEEC_REQUIRE(r.converged, FEEC_E_NOT_CONVERGED, "Dirac source solve did not converge");
  });
}

}  // extern "C"
