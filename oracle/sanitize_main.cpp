// Driver of the oracle under AddressSanitizer and UndefinedBehaviorSanitizer
// (test infrastructure: built and run by tests/test_oracle_sanitizers.py, never
// part of the product).  It reads a scene written by that test and exercises
// every entry point of xpbi_oracle.cpp on it: whole substeps in f32 and f64 with
// each start (plain, residual loop, training pairs, learned, given lambda0), and
// the unit entry points (kernels, SVD, return maps, HB bisection, grid hash,
// MLP forward) on the scene's own materials and on random inputs.
//
//   oracle_san INPUT   (exit 0: every call returned; a sanitizer finding aborts)
//
// INPUT: u32 N, u32 C, u32 hasWater, then the 256-B params record, eight 64-B
// material records, pos[4N], vel[4N], F[12N], aux[4N] (f32), id[N] (u32) and the
// 5122 MLP parameters (f32).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

typedef uint32_t u32;
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
struct xo_opts {
  double tol;
  int wsMode;
  const float* wsParams;
  void* feat;
  void* target;
  int itersDone;
  const void* lam0ById;
};
int xo_substep_ex_f32(const xo_state32*, const void*, const void*, int, int, int, int, double*, xo_opts*);
int xo_substep_ex_f64(const xo_state64*, const void*, const void*, int, int, int, int, double*, xo_opts*);
void xo_mlp_forward(const float*, int, const double*, double*);
void xo_kernel_f32(float, int, const float*, const float*, float*, float*, float*, float*);
void xo_kernel_f64(double, int, const double*, const double*, double*, double*, double*, double*);
void xo_svd3_f32(int, const float*, float*, float*, float*);
void xo_svd3_f64(int, const double*, double*, double*, double*);
void xo_return_map_f32(int, const void*, float, int, const float*, const float*, float*, float*, float*, float*);
void xo_return_map_f64(int, const void*, double, int, const double*, const double*, double*, double*, double*,
                       double*);
void xo_cell_index(const void*, int, const float*, u32*);
void xo_hb_return_map_f32(int, const void*, float, const float*, const float*, float*, float*);
void xo_hb_return_map_f64(int, const void*, double, const double*, const double*, double*, double*);
}

template <class T>
struct Run {
  u32 N, C;
  std::vector<T> pos, velA, velB, F, L, aux, Sp;
  std::vector<u32> cellTab, sortedIdx, id;
  Run(u32 n, u32 c, const std::vector<float>& p, const std::vector<float>& v, const std::vector<float>& f,
      const std::vector<float>& a, const std::vector<u32>& ids)
      : N(n), C(c), pos(p.begin(), p.end()), velA(v.begin(), v.end()), velB(4ull * n), F(f.begin(), f.end()),
        L(12ull * n), aux(a.begin(), a.end()), Sp(12ull * n), cellTab(c), sortedIdx(n), id(ids) {}
  template <class S>
  S st() {
    return S{N, C, pos.data(), velA.data(), velB.data(), F.data(), L.data(), aux.data(), Sp.data(),
             cellTab.data(), sortedIdx.data(), id.data()};
  }
};

static void need(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "oracle_san: %s\n", what);
    std::exit(2);
  }
}

template <class T, class S, class Fn>
static void substeps(const char* label, Run<T> base, const uint8_t* P, const uint8_t* M, int hasWater,
                     const std::vector<float>& mlp, Fn&& call) {
  const u32 N = base.N;
  std::vector<T> feat(16ull * N), tgt(2ull * N);
  std::vector<T> lam0(N);
  for (u32 i = 0; i < N; i++) lam0[i] = T(1e-3) * T(i % 7);
  std::vector<double> resid(2 * 64);
  struct Case {
    const char* name;
    int iters, mode;
    double tol;
  } cases[] = {{"plain", 6, 0, 0.0}, {"residual", 40, 0, 1e-3}, {"pairs", 20, 2, 1e-3},
               {"learned", 20, 1, 1e-3}, {"given", 20, 3, 1e-3}, {"partial", 4, 0, 0.0}};
  for (const Case& c : cases) {
    Run<T> r = base;
    S s = r.template st<S>();
    xo_opts o{c.tol, c.mode, mlp.data(), feat.data(), tgt.data(), 0, lam0.data()};
    for (int k = 0; k < 2; k++) {
      // "partial": stop after computeL, then after the solve; the others: full substeps
      const int stop = std::strcmp(c.name, "partial") == 0 ? 3 + k : 7;
      need(call(&s, P, M, c.iters, hasWater, stop, 2, resid.data(), &o) == 0, "substep rejected the state");
    }
    double sum = 0;
    for (u32 i = 0; i < 4 * N; i++) sum += (double)r.pos[i];
    std::printf("%s %s: %d iterations, position sum %.6e\n", label, c.name, o.itersDone, sum);
  }
}

int main(int argc, char** argv) {
  need(argc == 2, "usage: oracle_san INPUT");
  FILE* f = std::fopen(argv[1], "rb");
  need(f != nullptr, "cannot open the input");
  u32 hdr[3];
  need(std::fread(hdr, 4, 3, f) == 3, "short header");
  const u32 N = hdr[0], C = hdr[1], hasWater = hdr[2];
  std::vector<uint8_t> P(256), M(512);
  std::vector<float> pos(4ull * N), vel(4ull * N), F(12ull * N), aux(4ull * N), mlp(5122);
  std::vector<u32> id(N);
  need(std::fread(P.data(), 1, 256, f) == 256 && std::fread(M.data(), 1, 512, f) == 512, "short records");
  need(std::fread(pos.data(), 4, pos.size(), f) == pos.size() && std::fread(vel.data(), 4, vel.size(), f) == vel.size() &&
           std::fread(F.data(), 4, F.size(), f) == F.size() && std::fread(aux.data(), 4, aux.size(), f) == aux.size() &&
           std::fread(id.data(), 4, id.size(), f) == id.size() && std::fread(mlp.data(), 4, mlp.size(), f) == mlp.size(),
       "short arrays");
  std::fclose(f);

  Run<float> r32(N, C, pos, vel, F, aux, id);
  Run<double> r64(N, C, pos, vel, F, aux, id);
  substeps<float, xo_state32>("f32", r32, P.data(), M.data(), (int)hasWater, mlp, xo_substep_ex_f32);
  substeps<double, xo_state64>("f64", r64, P.data(), M.data(), (int)hasWater, mlp, xo_substep_ex_f64);

  // unit entry points on random inputs and on each material record
  const int n = 257;
  std::mt19937 g(7);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  std::vector<float> r(n), d(3 * n), W(n), gW(3 * n), p6(n), sp(3 * n);
  for (int i = 0; i < n; i++) {
    r[i] = 0.03f * std::fabs(u(g));
    for (int k = 0; k < 3; k++) d[3 * i + k] = 0.02f * u(g);
  }
  xo_kernel_f32(0.01f, n, r.data(), d.data(), W.data(), gW.data(), p6.data(), sp.data());
  std::vector<double> rd(r.begin(), r.end()), dd(d.begin(), d.end()), Wd(n), gWd(3 * n), p6d(n), spd(3 * n);
  xo_kernel_f64(0.01, n, rd.data(), dd.data(), Wd.data(), gWd.data(), p6d.data(), spd.data());
  std::vector<float> Fm(9 * n), U(9 * n), sig(3 * n), V(9 * n);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 9; k++) Fm[9 * i + k] = (k % 4 == 0 ? 1.0f : 0.0f) + 0.3f * u(g);
  for (int k = 0; k < 9; k++) Fm[k] = 0.0f;  // a zero matrix and a rank-one matrix among them
  for (int k = 0; k < 9; k++) Fm[9 + k] = (k < 3) ? 1.0f : 0.0f;
  xo_svd3_f32(n, Fm.data(), U.data(), sig.data(), V.data());
  std::vector<double> Fd(Fm.begin(), Fm.end()), Ud(9 * n), sigd(3 * n), Vd(9 * n);
  xo_svd3_f64(n, Fd.data(), Ud.data(), sigd.data(), Vd.data());
  std::vector<float> s3(3 * n), st(n), so(3 * n), sto(n), psi(n), dpsi(3 * n), hint(n), tout(n);
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) s3[3 * i + k] = std::exp(0.2f * u(g));
    st[i] = 0.01f * u(g);
    hint[i] = (i % 3 == 0) ? -1.0f : 0.5f + 0.5f * u(g);
  }
  std::vector<double> s3d(s3.begin(), s3.end()), std_(st.begin(), st.end()), sod(3 * n), stod(n), psid(n),
      dpsid(3 * n), hintd(hint.begin(), hint.end()), toutd(n);
  for (int m = 0; m < 8; m++) {
    const uint8_t* mat = M.data() + 64 * m;
    for (int harden = 0; harden < 2; harden++) {
      xo_return_map_f32(n, mat, 1.0f / 6000.0f, harden, s3.data(), st.data(), so.data(), sto.data(), psi.data(),
                        dpsi.data());
      xo_return_map_f64(n, mat, 1.0 / 6000.0, harden, s3d.data(), std_.data(), sod.data(), stod.data(), psid.data(),
                        dpsid.data());
    }
    u32 kind;
    std::memcpy(&kind, mat + 12, 4);
    if (kind == 4u) {  // Herschel-Bulkley: the bisection with and without a bracket hint
      xo_hb_return_map_f32(n, mat, 1.0f / 6000.0f, s3.data(), hint.data(), so.data(), tout.data());
      xo_hb_return_map_f64(n, mat, 1.0 / 6000.0, s3d.data(), hintd.data(), sod.data(), toutd.data());
    }
  }
  std::vector<u32> cell(N);
  std::vector<float> p3(3ull * N);
  for (u32 i = 0; i < N; i++)
    for (int k = 0; k < 3; k++) p3[3 * i + k] = pos[4ull * i + k];
  xo_cell_index(P.data(), (int)N, p3.data(), cell.data());
  std::vector<double> x(12 * n), y(2 * n);
  for (int i = 0; i < 12 * n; i++) x[i] = u(g);
  xo_mlp_forward(mlp.data(), n, x.data(), y.data());
  std::printf("unit entry points done\n");
  return 0;
}
