// C entry points of the CPU oracle -- TEST INFRASTRUCTURE ONLY (loaded by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg through
// ctypes; never linked into the product library).
//
// How each restated function is pinned (tests/test_oracle_*.py):
//   kernelW / kernelGradW     closed form at sample radii, integral of W over R^3
//                             equals 1 (radial quadrature), central finite
//                             difference of W equals kernelGradW, antisymmetry
//   kernelPoly6 / SpikyGrad   closed forms; spiky grad is the derivative of the
//                             spiky kernel (finite difference)
//   cellCoord / cellIndex     brute force floor/clamp on random points, including
//                             points outside the grid and NaN
//   counting sort             sortedIdx equals numpy's stable argsort of the keys;
//                             cellTab equals the exclusive prefix sum of bincount
//   svd3                      U diag(sig) V^T reproduces F; U, V orthonormal with
//                             det +1; sig(3) < 0 exactly when det F < 0; |sig|
//                             equals LAPACK (numpy.linalg.svd) singular values
//   stvkPsi / stvkDPsi        Psi(I) = 0, closed form on diagonal input, finite
//                             difference of Psi equals dPsi
//   rmDruckerPrager           projected state lies on the cone (delta-gamma = 0),
//                             elastic states unchanged, tension -> tip, trace of
//                             Hencky strain unchanged by the cone projection
//   rmVonMises                |dev eps| = sigY/(2 mu) after projection, trace kept
//   rmHerschelBulkley         h = 1: matches the closed form; general h: the root
//                             of g(s) lies inside the final 2^-18 bracket and the
//                             bisection agrees with a 200-step double bisection
//   rmSnowClamp / snow state  clamp bounds and the Jp update identity
//   rmNACC                    yield value y <= 1e-4 after a case-3 projection,
//                             tip formulas for cases 1/2
//   solveA/solveB, xsph       (test_oracle_substep.py) a rest lattice has C = 0,
//                             zero Sp and stays exactly at rest; a substep away
//                             from walls keeps the total momentum (constraint
//                             gather, XSPH and contacts are pairwise); dropping
//                             selfDv, flipping the gather's gradient or XSPH's
//                             velocity difference each fails one of these
//   substep                   free fall: every particle has the same velocity,
//                             the sum of the dt g steps, and F = I (f32 and f64);
//                             a rigid translation gets no impulse at all
//   pbfA/pbfB                 water below rho0 (checked with the oracle's own
//                             poly6) gets lambda_pbf = 0 and, without s_corr,
//                             stays at rest
//   computeL + velocity grad (test_oracle_pins.py) a linear velocity field v = G x gives the
//                             corrected gradient G P, P the projection onto the right
//                             singular vectors of A (numpy/LAPACK) above 0.2, including a
//                             particle with two directions dropped
//   solveA + solveB           the first damped XPBD step from lambda = 0 for one active
//                             constraint equals -(C + gamma dC/dv.v) / ((1 + gamma)
//                             sum |dC/dv_k|^2 / (m_k dt^2) + 1 / (V0 dt^2)) with dC/dv_k
//                             from central differences of C, and every velocity moves by
//                             dlambda dC/dv_k / (m_k dt^2)
//   applyBoundary             floor: v_n = (bound - x) / dt, tangential speed cut by
//                             min(mu |dv_n|, |v_t|); sphere: outward relative speed
//                             min((R - d) / dt, 0.6), tangential motion kept
//   contacts                  an isolated pair below 0.75 h: push (0.75 h - d) shared
//                             as m_b / (m_a + m_b), momentum kept
//   pbfA / pbfB               lambda_j (|J_j|^2 + eps) = -C_j with J the Jacobian of the
//                             density constraints (eps 100 and 0); correction J^T lambda;
//                             s_corr(q H) = -k ((1 - q^2) / 0.96)^12 at four distances
//   xsph                      isolated pairs: c V_b (v_b - v_i) W(r), 10 c for water, no
//                             solid-water smoothing
//   planted mistakes          test_oracle_mutations.py: 15 plausible slips (pinv
//                             tolerance, V_b in A, alpha~, (1 + gamma), the damping sign,
//                             selfDv, gather sign, contact weight, friction, wall speed,
//                             sphere cap, PBF eps, PBF neighbour lambda, s_corr exponent,
//                             XSPH sign) each fail one of these pins
// Parity of this oracle with the reference WGSL *as executed by a WebGPU stack*
// is unpinned: no WebGPU implementation exists in this environment.
#include "xpbi_oracle.hpp"

using namespace xo;

extern "C" {
struct xo_state32 {
  u32 N, C;
  float *pos, *velA, *velB, *F, *L, *aux, *Sp;
  u32 *cellTab, *sortedIdx, *id;
};
struct xo_state64 {
  u32 N, C;
  double *pos, *velA, *velB, *F, *L, *aux, *Sp;
  u32 *cellTab, *sortedIdx, *id;
};
}  // extern "C"

// learned warm start / residual loop options of a substep (no reference counterpart)
struct xo_opts {
  double tol;             // > 0: iterations is a cap, stop at residual <= tol
  int wsMode;             // 0 plain, 1 learned warm start, 2 collect training pairs
  const float* wsParams;  // canonical MLP parameters (mode 1)
  void* feat;             // N x 16 feature rows (modes 1, 2), T
  void* target;           // N x 2 targets (mode 2), T
  int itersDone;          // out: Jacobi iterations applied
  const void* lam0ById;   // mode 3: lambda0 per persistent particle id, T
};

template <class T, class ST>
static int run_substep(const ST* st, const void* params, const void* mats, int iters, int hasWater, int stop,
                       int nthreads, double* resid, xo_opts* o) {
  Sim<T> sim;
  ParamRec pr;
  std::memcpy(&pr, params, sizeof(pr));
  sim.P = toPrm<T>(pr);
  const MatRec* mr = (const MatRec*)mats;
  for (int k = 0; k < 8; k++) sim.M[k] = toMat<T>(mr[k]);
  if (st->N != pr.N || st->C != pr.numCells) return -1;
  sim.S.N = st->N; sim.S.C = st->C;
  sim.S.pos = st->pos; sim.S.velA = st->velA; sim.S.velB = st->velB; sim.S.F = st->F; sim.S.L = st->L;
  sim.S.aux = st->aux; sim.S.Sp = st->Sp; sim.S.cellTab = st->cellTab; sim.S.sortedIdx = st->sortedIdx;
  sim.S.id = st->id;
  sim.nthreads = nthreads > 0 ? nthreads : 1;
  sim.resid = resid;
  if (o) {
    if ((o->wsMode == 1 && !o->wsParams) || (o->wsMode && !o->feat) || (o->wsMode == 2 && !o->target) ||
        (o->wsMode == 3 && !o->lam0ById))
      return -2;
    sim.tol = o->tol;
    sim.wsMode = o->wsMode;
    sim.wsParams = o->wsParams;
    sim.feat = (T*)o->feat;
    sim.target = (T*)o->target;
    sim.lam0ById = (const T*)o->lam0ById;
  }
  sim.substep(iters, hasWater != 0, stop);
  if (o) o->itersDone = sim.itersDone;
  return 0;
}

extern "C" int xo_substep_f32(const xo_state32* st, const void* params, const void* mats, int iters, int hasWater, int stop,
                   int nthreads, double* resid) {
  return run_substep<float>(st, params, mats, iters, hasWater, stop, nthreads, resid, nullptr);
}
extern "C" int xo_substep_f64(const xo_state64* st, const void* params, const void* mats, int iters, int hasWater, int stop,
                   int nthreads, double* resid) {
  return run_substep<double>(st, params, mats, iters, hasWater, stop, nthreads, resid, nullptr);
}
extern "C" int xo_substep_ex_f32(const xo_state32* st, const void* params, const void* mats, int iters, int hasWater,
                                 int stop, int nthreads, double* resid, xo_opts* o) {
  return run_substep<float>(st, params, mats, iters, hasWater, stop, nthreads, resid, o);
}
extern "C" int xo_substep_ex_f64(const xo_state64* st, const void* params, const void* mats, int iters, int hasWater,
                                 int stop, int nthreads, double* resid, xo_opts* o) {
  return run_substep<double>(st, params, mats, iters, hasWater, stop, nthreads, resid, o);
}
// the C++ restatement of the MLP forward on n feature rows of 12 (pinned against the numpy oracle)
extern "C" void xo_mlp_forward(const float* params, int n, const double* x, double* y) {
  for (int i = 0; i < n; i++) mlpForward<double>(params, x + 12 * (size_t)i, y + 2 * (size_t)i);
}

// ---- unit entry points (batched over n inputs) ----
static Prm<float> prmF(float h) {
  ParamRec r{};
  r.h = h; r.support = 2 * h; r.cellSize = 2 * h;
  return toPrm<float>(r);
}
static Prm<double> prmD(double h) {
  Prm<double> p = toPrm<double>(ParamRec{});
  p.h = h; p.support = 2 * h;
  return p;
}

extern "C" void xo_kernel_f32(float h, int n, const float* r, const float* dxv, float* W, float* gradW, float* poly6,
                   float* spiky) {
  Prm<float> P = prmF(h);
  for (int i = 0; i < n; i++) {
    W[i] = kernelW(P, r[i]);
    poly6[i] = kernelPoly6(P, r[i]);
    v3<float> d = V(dxv[3 * i], dxv[3 * i + 1], dxv[3 * i + 2]);
    v3<float> g = kernelGradW(P, d);
    v3<float> s = kernelSpikyGrad(P, d);
    for (int k = 0; k < 3; k++) { gradW[3 * i + k] = g[k]; spiky[3 * i + k] = s[k]; }
  }
}
extern "C" void xo_kernel_f64(double h, int n, const double* r, const double* dxv, double* W, double* gradW, double* poly6,
                   double* spiky) {
  Prm<double> P = prmD(h);
  for (int i = 0; i < n; i++) {
    W[i] = kernelW(P, r[i]);
    poly6[i] = kernelPoly6(P, r[i]);
    v3<double> d = V(dxv[3 * i], dxv[3 * i + 1], dxv[3 * i + 2]);
    v3<double> g = kernelGradW(P, d);
    v3<double> s = kernelSpikyGrad(P, d);
    for (int k = 0; k < 3; k++) { gradW[3 * i + k] = g[k]; spiky[3 * i + k] = s[k]; }
  }
}

// F, U, V: 9 scalars per matrix, column-major (F[3*col+row])
template <class T>
static void svd_batch(int n, const T* F, T* U, T* sig, T* Vv) {
  for (int i = 0; i < n; i++) {
    m3<T> m;
    for (int c = 0; c < 3; c++) m.c[c] = V(F[9 * i + 3 * c], F[9 * i + 3 * c + 1], F[9 * i + 3 * c + 2]);
    SVD<T> s = svd3(m);
    for (int c = 0; c < 3; c++)
      for (int r = 0; r < 3; r++) { U[9 * i + 3 * c + r] = s.U.c[c][r]; Vv[9 * i + 3 * c + r] = s.V.c[c][r]; }
    for (int k = 0; k < 3; k++) sig[3 * i + k] = s.sig[k];
  }
}
extern "C" void xo_svd3_f32(int n, const float* F, float* U, float* sig, float* Vv) { svd_batch<float>(n, F, U, sig, Vv); }
extern "C" void xo_svd3_f64(int n, const double* F, double* U, double* sig, double* Vv) { svd_batch<double>(n, F, U, sig, Vv); }

// return map + StVK on n singular-value triples; mat is one 64-B record;
// harden != 0 applies hardenedMat(m, state) first, as solveA/updateF do.
template <class T>
static void rm_batch(int n, const void* mat, T dt, int harden, const T* sig, const T* state, T* sigOut, T* stateOut,
                     T* psi, T* dpsi) {
  MatRec mr;
  std::memcpy(&mr, mat, sizeof(mr));
  Mat<T> m0 = toMat<T>(mr);
  for (int i = 0; i < n; i++) {
    Mat<T> m = harden ? hardenedMat(m0, state[i]) : m0;
    v3<T> s = V(sig[3 * i], sig[3 * i + 1], sig[3 * i + 2]);
    T so;
    v3<T> r = returnMap(s, m, state[i], dt, &so);
    for (int k = 0; k < 3; k++) sigOut[3 * i + k] = r[k];
    stateOut[i] = so;
    psi[i] = stvkPsi(r, m.mu, m.lam);
    v3<T> d = stvkDPsi(r, m.mu, m.lam);
    for (int k = 0; k < 3; k++) dpsi[3 * i + k] = d[k];
  }
}
extern "C" void xo_return_map_f32(int n, const void* mat, float dt, int harden, const float* sig, const float* state,
                       float* sigOut, float* stateOut, float* psi, float* dpsi) {
  rm_batch<float>(n, mat, dt, harden, sig, state, sigOut, stateOut, psi, dpsi);
}
extern "C" void xo_return_map_f64(int n, const void* mat, double dt, int harden, const double* sig, const double* state,
                       double* sigOut, double* stateOut, double* psi, double* dpsi) {
  rm_batch<double>(n, mat, dt, harden, sig, state, sigOut, stateOut, psi, dpsi);
}

// grid hash of n points (x,y,z stride 3) with a 256-B param record
extern "C" void xo_cell_index(const void* params, int n, const float* p, u32* out) {
  ParamRec pr;
  std::memcpy(&pr, params, sizeof(pr));
  Prm<float> P = toPrm<float>(pr);
  for (int i = 0; i < n; i++) out[i] = cellIndex(P, cellCoord(P, p[3 * i], p[3 * i + 1], p[3 * i + 2]));
}

// the HB return map with a bracket hint (learned warm start): sigOut, root position tOut
template <class T>
static void hb_batch(int n, const void* mat, T dt, const T* sig, const T* hint, T* sigOut, T* tOut) {
  const Mat<T> m = toMat<T>(*(const MatRec*)mat);
  for (int i = 0; i < n; i++) {
    v3<T> s = rmHerschelBulkley(V(sig[3 * i], sig[3 * i + 1], sig[3 * i + 2]), m, dt, hint[i], &tOut[i]);
    sigOut[3 * i] = s.x; sigOut[3 * i + 1] = s.y; sigOut[3 * i + 2] = s.z;
  }
}
extern "C" void xo_hb_return_map_f32(int n, const void* mat, float dt, const float* sig, const float* hint,
                                     float* sigOut, float* tOut) {
  hb_batch<float>(n, mat, dt, sig, hint, sigOut, tOut);
}
extern "C" void xo_hb_return_map_f64(int n, const void* mat, double dt, const double* sig, const double* hint,
                                     double* sigOut, double* tOut) {
  hb_batch<double>(n, mat, dt, sig, hint, sigOut, tOut);
}
