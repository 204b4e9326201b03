// C ABI of libxpbi (declared in include/xpbi.h).  Host orchestration of the
// substep: grid set-up and state upload (web/js/sim.js:10-82), material and
// parameter records (sim.js:84-99, :159-186), the dispatch schedule
// (sim.js:220-272), per-launch profiling (sim.js:188-215) and buffer readback
// (web/js/main.js:289-313).
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "xpbi_internal.hpp"

using namespace xpbi;

namespace {

thread_local std::string g_create_err;

int fail(Sim* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  else g_create_err = msg;
  return code;
}
int hip_fail(Sim* s, hipError_t e, const char* what) {
  return fail(s, XPBI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T> hipError_t dalloc(Sim* s, T** p, size_t count) {
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = 16;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return e;
  s->allocs.push_back(q);
  *p = (T*)q;
  return hipSuccess;
}

void drop_graph(Sim* s) {
  if (s->graphExec) hipGraphExecDestroy(s->graphExec);
  if (s->graph) hipGraphDestroy(s->graph);
  s->graphExec = nullptr;
  s->graph = nullptr;
}

void free_all(Sim* s) {
  if (s->stream) hipStreamSynchronize(s->stream);
  drop_graph(s);
  for (void* p : s->allocs) hipFree(p);
  s->allocs.clear();
  if (s->ctlHost) hipHostFree(s->ctlHost);
  s->ctlHost = nullptr;
  if (s->mirror) hipHostFree(s->mirror);
  s->mirror = nullptr;
  for (hipEvent_t e : s->events) hipEventDestroy(e);
  s->events.clear();
  if (s->stream) hipStreamDestroy(s->stream);
  s->stream = nullptr;
}

// derived constants, formed in f32 with no contraction so that they equal the
// per-pair expressions of common.js evaluated in f32
#pragma clang fp contract(off)
void build_kp(Scene& sc) {
  const xpbi_params& p = sc.params;
  KP k{};
  k.ox = p.gridOrigin[0]; k.oy = p.gridOrigin[1]; k.oz = p.gridOrigin[2]; k.cell = p.cellSize;
  k.gx = p.gridDims[0]; k.gy = p.gridDims[1]; k.gz = p.gridDims[2]; k.C = p.numCells;
  k.dminx = p.domainMin[0]; k.dminy = p.domainMin[1]; k.dminz = p.domainMin[2]; k.dt = p.dt;
  k.dmaxx = p.domainMax[0]; k.dmaxy = p.domainMax[1]; k.dmaxz = p.domainMax[2]; k.h = p.h;
  k.grx = p.gravity[0]; k.gry = p.gravity[1]; k.grz = p.gravity[2]; k.support = p.support;
  k.cpx = p.colliderPos[0]; k.cpy = p.colliderPos[1]; k.cpz = p.colliderPos[2]; k.cR = p.colliderR;
  k.cvx = p.colliderVel[0]; k.cvy = p.colliderVel[1]; k.cvz = p.colliderVel[2]; k.cActive = p.colliderActive;
  k.N = p.N; k.V0 = p.V0; k.xsphC = p.xsphC; k.omega = p.omega;
  k.pbfEps = p.pbfEps; k.sCorrK = p.sCorrK; k.contactOmega = p.contactOmega; k.kDamp = p.kDamp;
  const float h = p.h, H = p.support;
  k.supp2 = H * H;
  k.sigW = 21.0f / ((float)(16.0 * 3.14159265) * h * h * h);
  k.poly6C = 315.0f / ((float)(64.0 * 3.14159265) * std::pow(H, 9.0f));
  k.spikyC = -45.0f / (3.14159265f * std::pow(H, 6.0f));
  {
    float r = 0.2f * H;
    float x = H * H - r * r;
    k.w0 = (r >= H) ? 0.0f : k.poly6C * x * x * x;
    float x0 = H * H - 0.0f * 0.0f;
    k.poly6_0 = k.poly6C * x0 * x0 * x0;
  }
  k.alphaT = (1.0f / p.V0) / (p.dt * p.dt);
  k.vmax = 0.45f * H / p.dt;
  k.dvCap = 0.2f * H / p.dt;
  k.minD = 0.75f * h;
  k.gxy = (u32)(p.gridDims[0] * p.gridDims[1]);
  k.invH = 1.0f / h;
  k.halfInvH = 0.5f / h;
  k.gradC = -5.0f * k.sigW / (h * h);
  k.invW0 = 1.0f / std::max(k.w0, 1e-12f);
  k.cellBase = sc.cellBase;
  k.pBase = sc.pBase;
  k.pEnd = sc.pBase + sc.N;
  sc.kp = k;
}
#pragma clang fp contract(on)

int check_launch(Sim* s, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(s, e, what);
  return XPBI_OK;
}

int run_substep(Sim* s, int iterations, int has_water, int stop) {
  for (const Scene& sc : s->sc) {
    if (!sc.haveParams) return fail(s, XPBI_ERR_STATE, "substep before xpbi_write_params");
    if (!sc.haveMats) return fail(s, XPBI_ERR_STATE, "substep before xpbi_write_materials");
  }
  if (iterations < 0) return fail(s, XPBI_ERR_ARG, "iterations < 0");
  if (stop < XPBI_STAGE_SORT || stop > XPBI_STAGE_FULL) return fail(s, XPBI_ERR_ARG, "bad stage");
  hipSetDevice(s->dev);
  if (s->N == 0) return XPBI_OK;
  launch_sort(s);
  s->sorted = true;
  if (stop == XPBI_STAGE_SORT) return check_launch(s, "sort");
  launch_reorder(s);
  if (stop == XPBI_STAGE_REORDER) return check_launch(s, "reorder");
  launch_solve_stages(s, iterations, has_water != 0, stop);
  return check_launch(s, "substep");
}

size_t field_bytes(const Sim* s, int field) {
  switch (field) {
    case XPBI_FIELD_POS:
    case XPBI_FIELD_VEL:
    case XPBI_FIELD_AUX: return 16ull * s->N;
    case XPBI_FIELD_F:
    case XPBI_FIELD_L:
    case XPBI_FIELD_SP: return 48ull * s->N;
    case XPBI_FIELD_CELLSTART: return 4ull * s->C;
    case XPBI_FIELD_PERM:
    case XPBI_FIELD_ID: return 4ull * s->N;
    case XPBI_FIELD_MATS: return sizeof(DevMat) * 8ull * s->sc.size();
    default: return 0;
  }
}

// planes (a, b, c) <-> reference 3-column vec4 layout (pad 0, or w from `w4`)
void planes_to_ref(u32 N, const float* a, const float* b, const float* c, const float* w4, float* out) {
  for (u32 i = 0; i < N; i++) {
    const float* A = a + 4ull * i;
    const float* B = b + 4ull * i;
    float m[9] = {A[0], A[1], A[2], A[3], B[0], B[1], B[2], B[3], c[i]};
    float* o = out + 12ull * i;
    for (int col = 0; col < 3; col++) {
      o[4 * col + 0] = m[3 * col + 0];
      o[4 * col + 1] = m[3 * col + 1];
      o[4 * col + 2] = m[3 * col + 2];
      o[4 * col + 3] = w4 ? w4[4ull * i + col] : 0.0f;
    }
  }
}
void ref_to_planes(u32 N, const float* in, float* a, float* b, float* c) {
  for (u32 i = 0; i < N; i++) {
    const float* o = in + 12ull * i;
    float m[9];
    for (int col = 0; col < 3; col++)
      for (int r = 0; r < 3; r++) m[3 * col + r] = o[4 * col + r];
    for (int k = 0; k < 4; k++) { a[4ull * i + k] = m[k]; b[4ull * i + k] = m[4 + k]; }
    c[i] = m[8];
  }
}

int d2h(Sim* s, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return XPBI_OK;
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? XPBI_OK : hip_fail(s, e, "hipMemcpy D2H");
}
int h2d(Sim* s, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return XPBI_OK;
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? XPBI_OK : hip_fail(s, e, "hipMemcpy H2D");
}

}  // namespace

extern "C" {

int xpbi_create(const xpbi_scene_desc* d, xpbi_sim** out) { return xpbi_create_batch(d, 1, out); }

int xpbi_create_batch(const xpbi_scene_desc* descs, int count, xpbi_sim** out) {
  if (!out) return fail(nullptr, XPBI_ERR_ARG, "out is null");
  *out = nullptr;
  if (!descs) return fail(nullptr, XPBI_ERR_ARG, "desc is null");
  if (count < 1 || count > 65535) return fail(nullptr, XPBI_ERR_ARG, "scene count must be 1..65535");
  std::vector<Scene> scs(count);
  u64 Nt = 0, Ct = 0, nbt = 0;
  for (int q = 0; q < count; q++) {
    const xpbi_scene_desc* d = &descs[q];
    if (d->N > 0 && !d->positions) return fail(nullptr, XPBI_ERR_ARG, "positions is null");
    if (d->N > 0 && !d->tags && !d->matIds) return fail(nullptr, XPBI_ERR_ARG, "tags and matIds are both null");
    if (!(d->h > 0.0) || !std::isfinite(d->h)) return fail(nullptr, XPBI_ERR_ARG, "h must be > 0");
    if (d->device != descs[0].device) return fail(nullptr, XPBI_ERR_ARG, "the scenes of a batch share one device");
    // grid: sim.js:20-26 (double arithmetic, as in JavaScript)
    const double cs = 2.0 * d->h;
    const int pad = 2;
    double dimsD[3];
    for (int k = 0; k < 3; k++) {
      if (!(d->domainMax[k] >= d->domainMin[k])) return fail(nullptr, XPBI_ERR_ARG, "domainMax < domainMin");
      dimsD[k] = std::ceil((d->domainMax[k] - d->domainMin[k]) / cs) + 2 * pad;
    }
    const double cells = dimsD[0] * dimsD[1] * dimsD[2];
    if (cells >= 4294967295.0 || dimsD[0] * dimsD[1] >= 2147483647.0)
      return fail(nullptr, XPBI_ERR_ARG, "grid too large");
    Scene& sc = scs[q];
    sc.N = d->N;
    sc.h = d->h;
    for (int k = 0; k < 3; k++) {
      sc.dims[k] = (int)dimsD[k];
      sc.origin[k] = (float)(d->domainMin[k] - pad * cs);
    }
    sc.cell = (float)cs;
    sc.C = (u32)cells;
    sc.pBase = (u32)Nt;
    sc.cellBase = (u32)Ct;
    Nt += sc.N;
    Ct += sc.C;
    nbt += (sc.N + 255) / 256;
  }
  // gathers address 16-B records by 32-bit byte offsets: at most 2^28 particles
  if (Nt >= (1ull << 28) || Ct >= 4294967295ull) return fail(nullptr, XPBI_ERR_ARG, "batch too large");

  xpbi_sim* h = new xpbi_sim();
  Sim* s = &h->s;
  s->dev = descs[0].device;
  s->sc = scs;
  s->N = (u32)Nt;
  s->C = (u32)Ct;
  s->nb = (u32)nbt;
  s->grid = s->nb;

  hipError_t e = hipSetDevice(s->dev);
  if (e != hipSuccess) { int r = hip_fail(nullptr, e, "hipSetDevice"); delete h; return r; }
  e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { int r = hip_fail(nullptr, e, "hipStreamCreate"); delete h; return r; }

  const u32 N = s->N, C = s->C;
  const u32 gN = s->nb;
  bool ok = true;
  auto A = [&](hipError_t r) { if (r != hipSuccess) ok = false; };
  for (int k = 0; k < 2; k++) {
    A(dalloc(s, &s->set[k].pos, N));
    A(dalloc(s, &s->set[k].Fa, N));
    A(dalloc(s, &s->set[k].Fb, N));
    A(dalloc(s, &s->set[k].Fc, N));
    A(dalloc(s, &s->set[k].aux, N));
    A(dalloc(s, &s->set[k].id, N));
    A(dalloc(s, &s->vel[k], N));
    A(dalloc(s, &s->keys[k], N));
    A(dalloc(s, &s->vals[k], N));
  }
  A(dalloc(s, &s->L.a, N));
  A(dalloc(s, &s->L.b, N));
  A(dalloc(s, &s->L.c, N));
  A(dalloc(s, &s->Sp.a, N));
  A(dalloc(s, &s->Sp.b, N));
  A(dalloc(s, &s->Sp.cx, N));
  A(dalloc(s, &s->selfDv, N));
  A(dalloc(s, &s->nbrLst, (size_t)s->nbrK * N));
  A(dalloc(s, &s->nbrHead, N));
  A(dalloc(s, &s->xc, N));
  A(dalloc(s, &s->cellStart, (size_t)C + 1));
  A(dalloc(s, &s->count, C));
  A(dalloc(s, &s->scanPart, scan_partials_needed(C)));
  A(dalloc(s, &s->dMats, 8ull * count));
  A(dalloc(s, &s->dKP, (size_t)count));
  A(dalloc(s, &s->vmaxSlot, 2ull * count));
  A(dalloc(s, &s->dTab, gN));
  A(dalloc(s, &s->residPart, 2ull * (gN > 0 ? gN : 1)));
  A(dalloc(s, &s->residOut, 2ull * kMaxResidIters));
  A(dalloc(s, &s->ctl, 1));
  A(hipHostMalloc((void**)&s->ctlHost, sizeof(SolveCtl), hipHostMallocDefault));
  A(hipHostMalloc((void**)&s->mirror, sizeof(CtlMirror), hipHostMallocCoherent | hipHostMallocMapped));
  if (s->mirror) std::memset(s->mirror, 0, sizeof(CtlMirror));
  if (!ok) {
    free_all(s);
    delete h;
    return fail(nullptr, XPBI_ERR_NOMEM, "device allocation failed");
  }
  s->perm = s->vals[0];
  s->sortedKeys = s->keys[0];

  // initial state: sim.js:52-81, scene after scene
  std::vector<float> pos(4ull * N), vel(4ull * N, 0.0f), aux(4ull * N, 0.0f), Fa(4ull * N, 0.0f),
      Fb(4ull * N, 0.0f), Fc(N, 1.0f);
  std::vector<u32> id(N), zeros((size_t)C + 1, 0u);
  std::vector<uint2> tab;
  tab.reserve(gN);
  for (int q = 0; q < count; q++) {
    const xpbi_scene_desc* d = &descs[q];
    const Scene& sc = s->sc[q];
    const float V0 = (float)(d->h * d->h * d->h);
    for (u32 j = 0; j < sc.N; j++) {
      const u64 i = sc.pBase + j;
      pos[4 * i + 0] = d->positions[3ull * j + 0];
      pos[4 * i + 1] = d->positions[3ull * j + 1];
      pos[4 * i + 2] = d->positions[3ull * j + 2];
      pos[4 * i + 3] = d->tags ? d->tags[j] : (float)d->matIds[j];
      if (d->velocities) {
        vel[4 * i + 0] = d->velocities[3ull * j + 0];
        vel[4 * i + 1] = d->velocities[3ull * j + 1];
        vel[4 * i + 2] = d->velocities[3ull * j + 2];
      }
      aux[4 * i + 0] = V0;
      aux[4 * i + 2] = d->states ? d->states[j] : 0.0f;
      Fa[4 * i + 0] = 1.0f;  // c0.x
      Fb[4 * i + 0] = 1.0f;  // c1.y
      id[i] = (u32)i;
    }
    for (u32 f = 0; f < sc.N; f += 256) tab.push_back(make_uint2((u32)q, sc.pBase + f));
  }
  int r = XPBI_OK;
  if (N > 0) {
    if (!r) r = h2d(s, s->set[0].pos, pos.data(), 16ull * N);
    if (!r) r = h2d(s, s->vel[0], vel.data(), 16ull * N);
    if (!r) r = h2d(s, s->set[0].aux, aux.data(), 16ull * N);
    if (!r) r = h2d(s, s->set[0].Fa, Fa.data(), 16ull * N);
    if (!r) r = h2d(s, s->set[0].Fb, Fb.data(), 16ull * N);
    if (!r) r = h2d(s, s->set[0].Fc, Fc.data(), 4ull * N);
    if (!r) r = h2d(s, s->set[0].id, id.data(), 4ull * N);
    if (!r) r = h2d(s, s->vals[0], id.data(), 4ull * N);
    if (!r) r = h2d(s, s->dTab, tab.data(), sizeof(uint2) * tab.size());
  }
  if (!r) r = h2d(s, s->cellStart, zeros.data(), 4ull * ((size_t)C + 1));
  if (r) {
    std::string m = s->err;
    free_all(s);
    delete h;
    return fail(nullptr, r, m);
  }
  s->events.resize(kMaxProfileMarks);
  for (auto& ev : s->events) hipEventCreate(&ev);
  s->labels.resize(kMaxProfileMarks);
  *out = h;
  return XPBI_OK;
}

void xpbi_destroy(xpbi_sim* h) {
  if (!h) return;
  hipSetDevice(h->s.dev);
  free_all(&h->s);
  delete h;
}

const char* xpbi_last_error(const xpbi_sim* h) { return h ? h->s.err.c_str() : g_create_err.c_str(); }

int xpbi_grid_info(const xpbi_sim* h, float origin[3], float* cell_size, int32_t dims[3], uint32_t* num_cells,
                   uint32_t* n_particles) {
  if (!h) return XPBI_ERR_ARG;
  return xpbi_scene_info(h, 0, origin, cell_size, dims, num_cells, n_particles, nullptr, nullptr);
}

int xpbi_scene_count(const xpbi_sim* h, int* count) {
  if (!h || !count) return XPBI_ERR_ARG;
  *count = (int)h->s.sc.size();
  return XPBI_OK;
}

int xpbi_scene_info(const xpbi_sim* h, int scene, float origin[3], float* cell_size, int32_t dims[3],
                    uint32_t* num_cells, uint32_t* n_particles, uint32_t* first_particle, uint32_t* first_cell) {
  if (!h) return XPBI_ERR_ARG;
  const Sim& s = h->s;
  if (scene < 0 || scene >= (int)s.sc.size()) return XPBI_ERR_ARG;
  const Scene& sc = s.sc[scene];
  for (int k = 0; k < 3; k++) {
    if (origin) origin[k] = sc.origin[k];
    if (dims) dims[k] = sc.dims[k];
  }
  if (cell_size) *cell_size = sc.cell;
  if (num_cells) *num_cells = sc.C;
  if (n_particles) *n_particles = sc.N;
  if (first_particle) *first_particle = sc.pBase;
  if (first_cell) *first_cell = sc.cellBase;
  return XPBI_OK;
}

int xpbi_write_materials(xpbi_sim* h, const xpbi_material* mats, int n) {
  return xpbi_write_scene_materials(h, 0, mats, n);
}

int xpbi_write_scene_materials(xpbi_sim* h, int scene, const xpbi_material* mats, int n) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (scene < 0 || scene >= (int)s->sc.size()) return fail(s, XPBI_ERR_ARG, "no such scene");
  if (n < 0 || n > 8) return fail(s, XPBI_ERR_ARG, "material count must be 0..8");
  if (n > 0 && !mats) return fail(s, XPBI_ERR_ARG, "mats is null");
  xpbi_material m[8];
  std::memset(m, 0, sizeof(m));
  for (int k = 0; k < n; k++) m[k] = mats[k];
  hipSetDevice(s->dev);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return hip_fail(s, e, "hipStreamSynchronize");
  int r = h2d(s, s->dMats + 8 * scene, m, sizeof(m));
  if (r) return r;
  Scene& sc = s->sc[scene];
  std::memcpy(sc.mats, m, sizeof(m));
  sc.haveMats = true;
  return XPBI_OK;
}

int xpbi_write_params(xpbi_sim* h, const xpbi_params* p) { return xpbi_write_scene_params(h, 0, p); }

int xpbi_write_scene_params(xpbi_sim* h, int scene, const xpbi_params* p) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (scene < 0 || scene >= (int)s->sc.size()) return fail(s, XPBI_ERR_ARG, "no such scene");
  if (!p) return fail(s, XPBI_ERR_ARG, "params is null");
  Scene& sc = s->sc[scene];
  for (int k = 0; k < 3; k++) {
    if (p->gridOrigin[k] != sc.origin[k] || p->gridDims[k] != sc.dims[k])
      return fail(s, XPBI_ERR_ARG, "params grid does not match the simulation grid");
  }
  if (p->cellSize != sc.cell || p->numCells != sc.C) return fail(s, XPBI_ERR_ARG, "params grid does not match");
  if (p->N != sc.N) return fail(s, XPBI_ERR_ARG, "params N does not match the particle count");
  if (!(p->dt > 0.0f) || !(p->h > 0.0f)) return fail(s, XPBI_ERR_ARG, "dt and h must be > 0");
  if (!(p->support > 0.0f) || p->support > p->cellSize)
    return fail(s, XPBI_ERR_ARG, "support must be in (0, cellSize]");
  hipSetDevice(s->dev);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return hip_fail(s, e, "hipStreamSynchronize");
  sc.params = *p;
  build_kp(sc);
  int r = h2d(s, s->dKP + scene, &sc.kp, sizeof(KP));
  if (r) return r;
  sc.haveParams = true;
  return XPBI_OK;
}

int xpbi_substep(xpbi_sim* h, int iterations, int has_water) {
  if (!h) return XPBI_ERR_ARG;
  return run_substep(&h->s, iterations, has_water, XPBI_STAGE_FULL);
}

int xpbi_substep_partial(xpbi_sim* h, int iterations, int has_water, int stage) {
  if (!h) return XPBI_ERR_ARG;
  return run_substep(&h->s, iterations, has_water, stage);
}

int xpbi_step(xpbi_sim* h, int substeps, int iterations, int has_water) {
  if (!h) return XPBI_ERR_ARG;
  if (substeps < 0) return fail(&h->s, XPBI_ERR_ARG, "substeps < 0");
  for (int k = 0; k < substeps; k++) {
    int r = run_substep(&h->s, iterations, has_water, XPBI_STAGE_FULL);
    if (r) return r;
  }
  return XPBI_OK;
}

int xpbi_sync(xpbi_sim* h) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  hipSetDevice(s->dev);
  hipError_t e = hipStreamSynchronize(s->stream);
  return e == hipSuccess ? XPBI_OK : hip_fail(s, e, "hipStreamSynchronize");
}

int xpbi_read(xpbi_sim* h, int field, void* dst, size_t bytes) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  size_t need = field_bytes(s, field);
  if (field < 0 || field >= XPBI_FIELD_COUNT) return fail(s, XPBI_ERR_ARG, "unknown field");
  if (bytes != need) return fail(s, XPBI_ERR_ARG, "byte count does not match the field size");
  if (need > 0 && !dst) return fail(s, XPBI_ERR_ARG, "dst is null");
  int r = xpbi_sync(h);
  if (r) return r;
  const u32 N = s->N;
  PSet& S = s->set[s->cur];
  switch (field) {
    case XPBI_FIELD_POS: return d2h(s, dst, S.pos, need);
    case XPBI_FIELD_VEL: return d2h(s, dst, s->vel[s->vcur], need);
    case XPBI_FIELD_AUX: return d2h(s, dst, S.aux, need);
    case XPBI_FIELD_CELLSTART: return d2h(s, dst, s->cellStart, need);
    case XPBI_FIELD_PERM: return d2h(s, dst, s->perm, need);
    case XPBI_FIELD_ID: return d2h(s, dst, S.id, need);
    case XPBI_FIELD_MATS: return d2h(s, dst, s->dMats, need);
    case XPBI_FIELD_F:
    case XPBI_FIELD_L:
    case XPBI_FIELD_SP: {
      std::vector<float> a(4ull * N), b(4ull * N), c(N), w;
      const float4 *pa, *pb;
      const float* pc = nullptr;
      if (field == XPBI_FIELD_F) { pa = S.Fa; pb = S.Fb; pc = S.Fc; }
      else if (field == XPBI_FIELD_L) { pa = s->L.a; pb = s->L.b; pc = s->L.c; }
      else { pa = s->Sp.a; pb = s->Sp.b; }
      if ((r = d2h(s, a.data(), pa, 16ull * N))) return r;
      if ((r = d2h(s, b.data(), pb, 16ull * N))) return r;
      if (pc) {
        if ((r = d2h(s, c.data(), pc, 4ull * N))) return r;
      } else {  // Sp's ninth float is the w lane of (x_i, c)
        std::vector<float> cx(4ull * N);
        if ((r = d2h(s, cx.data(), s->Sp.cx, 16ull * N))) return r;
        for (u32 i = 0; i < N; i++) c[i] = cx[4ull * i + 3];
      }
      if (field == XPBI_FIELD_SP) {
        w.resize(4ull * N);
        if ((r = d2h(s, w.data(), s->selfDv, 16ull * N))) return r;
      }
      planes_to_ref(N, a.data(), b.data(), c.data(), w.empty() ? nullptr : w.data(), (float*)dst);
      return XPBI_OK;
    }
  }
  return fail(s, XPBI_ERR_ARG, "unknown field");
}

int xpbi_write(xpbi_sim* h, int field, const void* src, size_t bytes) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (field < 0 || field >= XPBI_FIELD_COUNT) return fail(s, XPBI_ERR_ARG, "unknown field");
  size_t need = field_bytes(s, field);
  if (bytes != need) return fail(s, XPBI_ERR_ARG, "byte count does not match the field size");
  if (need > 0 && !src) return fail(s, XPBI_ERR_ARG, "src is null");
  int r = xpbi_sync(h);
  if (r) return r;
  const u32 N = s->N;
  PSet& S = s->set[s->cur];
  switch (field) {
    case XPBI_FIELD_POS: return h2d(s, S.pos, src, need);
    case XPBI_FIELD_VEL: return h2d(s, s->vel[s->vcur], src, need);
    case XPBI_FIELD_AUX: return h2d(s, S.aux, src, need);
    case XPBI_FIELD_ID: return h2d(s, S.id, src, need);
    case XPBI_FIELD_F: {
      std::vector<float> a(4ull * N), b(4ull * N), c(N);
      ref_to_planes(N, (const float*)src, a.data(), b.data(), c.data());
      if ((r = h2d(s, S.Fa, a.data(), 16ull * N))) return r;
      if ((r = h2d(s, S.Fb, b.data(), 16ull * N))) return r;
      return h2d(s, S.Fc, c.data(), 4ull * N);
    }
    default: return fail(s, XPBI_ERR_ARG, "field is read only");
  }
}

int xpbi_get_render_view(xpbi_sim* h, xpbi_render_view* out) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (!out) return fail(s, XPBI_ERR_ARG, "out is null");
  const PSet& S = s->set[s->cur];
  out->pos = S.pos;
  out->vel = s->vel[s->vcur];
  out->mats = s->dMats;
  out->stream = (void*)s->stream;
  out->N = s->N;
  out->scenes = (uint32_t)s->sc.size();
  out->device = s->dev;
  return XPBI_OK;
}

int xpbi_profile_substep(xpbi_sim* h, int iterations, int has_water, float* ms, int cap, int* count) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (cap < 0 || (cap > 0 && !ms)) return fail(s, XPBI_ERR_ARG, "bad ms/cap");
  int r = xpbi_sync(h);
  if (r) return r;
  s->profiling = true;
  s->nmarks = 0;
  hipEventRecord(s->events[0], s->stream);
  r = run_substep(s, iterations, has_water, XPBI_STAGE_FULL);
  s->profiling = false;
  if (r) return r;
  if ((r = xpbi_sync(h))) return r;
  int n = s->nmarks;
  for (int k = 0; k < n && k < cap; k++) {
    float t = 0.0f;
    hipEventElapsedTime(&t, s->events[k], s->events[k + 1]);
    ms[k] = t;
  }
  if (count) *count = n;
  return XPBI_OK;
}

const char* xpbi_profile_label(const xpbi_sim* h, int k) {
  if (!h || k < 0 || k >= h->s.nmarks) return "";
  return h->s.labels[k].c_str();
}

int xpbi_time_substeps(xpbi_sim* h, int substeps, int iterations, int has_water, float* ms) {
  if (!h || !ms) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  int r = xpbi_sync(h);
  if (r) return r;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, s->stream);
  for (int k = 0; k < substeps && !r; k++) r = run_substep(s, iterations, has_water, XPBI_STAGE_FULL);
  hipEventRecord(e1, s->stream);
  hipError_t e = hipEventSynchronize(e1);
  float t = 0.0f;
  hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (r) return r;
  if (e != hipSuccess) return hip_fail(s, e, "hipEventSynchronize");
  *ms = t;
  return XPBI_OK;
}

int xpbi_time_substeps_detail(xpbi_sim* h, int substeps, int iterations, int has_water, float* total_ms,
                              float* label_ms, int* label_count, int cap, int* nlabels) {
  if (!h || !total_ms || substeps < 0 || cap < 0 || (cap > 0 && (!label_ms || !label_count))) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  int r = xpbi_sync(h);
  if (r) return r;
  // one event per launch: at most 16 launches per substep outside the solve loop
  size_t need = (size_t)substeps * (size_t)(24 + 4 * (iterations > 0 ? iterations : 0)) + 2;
  while (s->events.size() < need) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) break;
    s->events.push_back(ev);
  }
  if (s->labels.size() < s->events.size()) s->labels.resize(s->events.size());
  s->profiling = true;
  s->nmarks = 0;
  hipEventRecord(s->events[0], s->stream);
  for (int k = 0; k < substeps && !r; k++) r = run_substep(s, iterations, has_water, XPBI_STAGE_FULL);
  s->profiling = false;
  if (r) return r;
  if ((r = xpbi_sync(h))) return r;
  float tot = 0.0f;
  hipEventElapsedTime(&tot, s->events[0], s->events[s->nmarks]);
  *total_ms = tot;
  std::vector<std::string> names;
  std::vector<double> sums;
  std::vector<int> counts;
  for (int k = 0; k < s->nmarks; k++) {
    float t = 0.0f;
    hipEventElapsedTime(&t, s->events[k], s->events[k + 1]);
    size_t j = 0;
    while (j < names.size() && names[j] != s->labels[k]) j++;
    if (j == names.size()) { names.push_back(s->labels[k]); sums.push_back(0.0); counts.push_back(0); }
    sums[j] += t;
    counts[j]++;
  }
  for (size_t j = 0; j < names.size(); j++) {
    s->labels[j] = names[j];
    if ((int)j < cap) { label_ms[j] = (float)sums[j]; label_count[j] = counts[j]; }
  }
  s->nmarks = (int)names.size();
  if (nlabels) *nlabels = (int)names.size();
  return XPBI_OK;
}

int xpbi_copy_bandwidth(xpbi_sim* h, size_t bytes, int reps, float* ms) {
  if (!h || !ms || reps < 1 || bytes < 16) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  int r = xpbi_sync(h);
  if (r) return r;
  u64 n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess) return fail(s, XPBI_ERR_NOMEM, "copy buffer");
  if (hipMalloc(&b, n * 16) != hipSuccess) { hipFree(a); return fail(s, XPBI_ERR_NOMEM, "copy buffer"); }
  hipMemsetAsync(a, 0, n * 16, s->stream);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  float t = 0.0f;
  for (int k = 0; k < reps; k++) {
    hipEventRecord(e0, s->stream);
    launch_copy(s, a, b, n);
    hipEventRecord(e1, s->stream);
  }
  hipError_t e = hipEventSynchronize(e1);
  hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return hip_fail(s, e, "copy");
  *ms = t;
  return XPBI_OK;
}

int xpbi_set_residual_monitor(xpbi_sim* h, int enabled) {
  if (!h) return XPBI_ERR_ARG;
  h->s.residOn = enabled != 0;
  return XPBI_OK;
}

int xpbi_read_residuals(xpbi_sim* h, double* out, int cap, int* count) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  int r = xpbi_sync(h);
  if (r) return r;
  int n = s->residIters;
  if (s->tol > 0.0f && n > 0) {
    // iterations past the stopping one left no entry
    SolveCtl c;
    if ((r = d2h(s, &c, s->ctl, sizeof(c)))) return r;
    if ((int)c.iters < n) n = (int)c.iters;
  }
  if (count) *count = n;
  int m = n < cap ? n : cap;
  if (m > 0) {
    if (!out) return fail(s, XPBI_ERR_ARG, "out is null");
    return d2h(s, out, s->residOut, 2ull * m * sizeof(double));
  }
  return XPBI_OK;
}

int xpbi_set_tolerance(xpbi_sim* h, float tol, int check_every) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (!(tol >= 0.0f) || check_every < 0) return fail(s, XPBI_ERR_ARG, "tol must be >= 0 and check_every >= 0");
  s->tol = tol;
  s->checkEvery = check_every;
  return XPBI_OK;
}

int xpbi_last_iterations(xpbi_sim* h, int* iterations, double* residual) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (!iterations) return fail(s, XPBI_ERR_ARG, "iterations is null");
  int r = xpbi_sync(h);
  if (r) return r;
  if (s->tol > 0.0f) {
    SolveCtl c;
    if ((r = d2h(s, &c, s->ctl, sizeof(c)))) return r;
    *iterations = (int)c.iters + s->warmIters;
    if (residual) *residual = c.c0sq > 0.0 ? sqrt(c.r2 / c.c0sq) : 0.0;
  } else {
    *iterations = s->lastIters + s->warmIters;
    if (residual) *residual = -1.0;
  }
  return XPBI_OK;
}

static int ensure_warmstart_buffers(Sim* s) {
  if (s->wsFeat) return XPBI_OK;
  hipSetDevice(s->dev);
  bool ok = true;
  auto A = [&](hipError_t r) { if (r != hipSuccess) ok = false; };
  const u32 N = s->N;
  A(dalloc(s, (uint8_t**)&s->wsPacked, mlp_packed_bytes()));
  A(dalloc(s, &s->wsFeat, 16ull * N));
  A(dalloc(s, &s->wsHb, N));
  A(dalloc(s, &s->wsTarget, 2ull * N));
  A(dalloc(s, &s->lamPrev, N));
  if (!ok) {
    s->wsFeat = nullptr;
    return fail(s, XPBI_ERR_NOMEM, "warm-start buffers: device allocation failed");
  }
  return XPBI_OK;
}

int xpbi_set_warmstart(xpbi_sim* h, const float* params, size_t n) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (!params) {
    s->wsMode = 0;
    return XPBI_OK;
  }
  if (n != (size_t)XPBI_MLP_PARAMS) return fail(s, XPBI_ERR_ARG, "params must hold XPBI_MLP_PARAMS floats");
  for (size_t k = 0; k < n; k++)
    if (!std::isfinite(params[k])) return fail(s, XPBI_ERR_ARG, "non-finite MLP parameter");
  int r = ensure_warmstart_buffers(s);
  if (r) return r;
  if ((r = xpbi_sync(h))) return r;
  std::vector<uint8_t> packed(mlp_packed_bytes());
  mlp_pack_host(params, packed.data());
  if ((r = h2d(s, s->wsPacked, packed.data(), packed.size()))) return r;
  s->wsMode = 1;
  return XPBI_OK;
}

int xpbi_set_lambda0(xpbi_sim* h, const float* lambda0, size_t n) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (!lambda0) {
    if (n != 0) return fail(s, XPBI_ERR_ARG, "lambda0 is null");
    if (s->wsMode == 3) s->wsMode = 0;
    return XPBI_OK;
  }
  if (n != (size_t)s->N) return fail(s, XPBI_ERR_ARG, "n must equal the particle count");
  int r = ensure_warmstart_buffers(s);
  if (r) return r;
  if (!s->lam0ById && dalloc(s, &s->lam0ById, (size_t)s->N) != hipSuccess) {
    s->lam0ById = nullptr;
    return fail(s, XPBI_ERR_NOMEM, "lambda0 buffer: device allocation failed");
  }
  if ((r = xpbi_sync(h))) return r;
  if ((r = h2d(s, s->lam0ById, lambda0, 4ull * n))) return r;
  s->wsMode = 3;
  return XPBI_OK;
}

int xpbi_collect_warmstart_pairs(xpbi_sim* h, int iterations, int has_water, float* x, float* t, float* w,
                                 size_t rows) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (rows != s->N) return fail(s, XPBI_ERR_ARG, "rows must equal the particle count");
  if (rows > 0 && (!x || !t)) return fail(s, XPBI_ERR_ARG, "x and t must not be null");
  int r = ensure_warmstart_buffers(s);
  if (r) return r;
  const int mode = s->wsMode;
  s->wsMode = 2;
  r = run_substep(s, iterations, has_water, XPBI_STAGE_FULL);
  s->wsMode = mode;
  if (r) return r;
  if ((r = xpbi_sync(h))) return r;
  if (rows == 0) return XPBI_OK;
  std::vector<float> f(16ull * rows);
  if ((r = d2h(s, f.data(), s->wsFeat, f.size() * sizeof(float)))) return r;
  if ((r = d2h(s, t, s->wsTarget, 2ull * rows * sizeof(float)))) return r;
  for (size_t i = 0; i < rows; i++) {
    memcpy(x + i * XPBI_MLP_INPUTS, &f[16 * i], XPBI_MLP_INPUTS * sizeof(float));
    if (w) w[i] = f[16 * i + 15] * f[16 * i + 15];
  }
  return XPBI_OK;
}

int xpbi_set_launch(xpbi_sim* h, int blocks) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (blocks < 0) return fail(s, XPBI_ERR_ARG, "blocks must be >= 0");
  u32 g = s->nb;
  if (blocks > 0) {
    g = ((u32)blocks + 7u) & ~7u;
    if (g >= s->nb) g = s->nb;
  }
  s->grid = g > 0 ? g : 1;
  s->persist = g < s->nb;
  return XPBI_OK;
}

int xpbi_get_launch(const xpbi_sim* h, int* blocks, int* work_entries) {
  if (!h) return XPBI_ERR_ARG;
  if (blocks) *blocks = h->s.persist ? (int)h->s.grid : 0;
  if (work_entries) *work_entries = (int)h->s.nb;
  return XPBI_OK;
}

int xpbi_set_neighbour_slots(xpbi_sim* h, int slots) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (slots < 0 || slots > (int)Sim::kMaxNbrSlots) return fail(s, XPBI_ERR_ARG, "slots must be in 0..1024");
  slots = (slots + 15) & ~15;  // the list stores slots in groups of 16
  if ((u32)slots == s->nbrK) return XPBI_OK;
  int r = xpbi_sync(h);
  if (r) return r;
  hipSetDevice(s->dev);
  drop_graph(s);  // a captured graph holds the old list
  for (auto it = s->allocs.begin(); it != s->allocs.end(); ++it) {
    if (*it == (void*)s->nbrLst) {
      s->allocs.erase(it);
      break;
    }
  }
  hipFree(s->nbrLst);
  s->nbrLst = nullptr;
  s->nbrK = 0;
  hipError_t e = dalloc(s, &s->nbrLst, (size_t)slots * s->N);
  if (e != hipSuccess) return hip_fail(s, e, "hipMalloc (neighbour list)");
  s->nbrK = (u32)slots;
  return XPBI_OK;
}

int xpbi_graph_step(xpbi_sim* h, int substeps, int iterations, int has_water) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (substeps < 0) return fail(s, XPBI_ERR_ARG, "substeps < 0");
  if (substeps == 0 || s->N == 0) return XPBI_OK;
  hipSetDevice(s->dev);
  // the graph bakes in the buffers, the loop shape and the launch shape
  float tol = s->tol;
  long long tolBits = 0;
  std::memcpy(&tolBits, &tol, sizeof(float));
  const long long key[8] = {1, iterations, has_water != 0, tolBits, s->wsMode, (long long)s->grid,
                            (long long)s->cur * 2 + s->vcur, s->residOn};
  if (!s->graphExec || std::memcmp(key, s->graphKey, sizeof(key)) != 0) {
    int r = xpbi_sync(h);
    if (r) return r;
    drop_graph(s);
    const int ce = s->checkEvery;
    const bool prof = s->profiling;
    s->checkEvery = 0;  // no host polling inside a graph
    s->profiling = false;
    hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      s->checkEvery = ce;
      s->profiling = prof;
      return hip_fail(s, e, "hipStreamBeginCapture");
    }
    // two substeps return the set and velocity ping-pong to where they started
    r = run_substep(s, iterations, has_water, XPBI_STAGE_FULL);
    if (!r) r = run_substep(s, iterations, has_water, XPBI_STAGE_FULL);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s->stream, &g);
    s->checkEvery = ce;
    s->profiling = prof;
    if (r) {
      if (g) hipGraphDestroy(g);
      return r;
    }
    if (e != hipSuccess) return hip_fail(s, e, "hipStreamEndCapture");
    s->graph = g;
    e = hipGraphInstantiate(&s->graphExec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(s);
      return hip_fail(s, e, "hipGraphInstantiate");
    }
    std::memcpy(s->graphKey, key, sizeof(key));
  }
  for (int k = 0; k + 1 < substeps; k += 2) {
    hipError_t e = hipGraphLaunch(s->graphExec, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "hipGraphLaunch");
  }
  if (substeps & 1) return run_substep(s, iterations, has_water, XPBI_STAGE_FULL);
  return XPBI_OK;
}

// ---- autotuner: the launch shape of the particle kernels, timed on this simulation
int xpbi_autotune(xpbi_sim* h, int iterations, int has_water, int substeps, int* best_blocks, float* best_ms) {
  if (!h) return XPBI_ERR_ARG;
  Sim* s = &h->s;
  if (substeps < 1) return fail(s, XPBI_ERR_ARG, "substeps must be >= 1");
  if (s->N == 0) return XPBI_OK;
  int r = xpbi_sync(h);
  if (r) return r;
  hipSetDevice(s->dev);
  int cus = 256;
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->dev);
  // the state the trials start from (and end on): current particle set and velocity
  const u32 N = s->N;
  PSet& S = s->set[s->cur];
  std::vector<void*> save;
  const std::pair<void*, size_t> fields[] = {{S.pos, 16ull * N}, {S.Fa, 16ull * N}, {S.Fb, 16ull * N},
                                             {S.Fc, 4ull * N},   {S.aux, 16ull * N}, {S.id, 4ull * N},
                                             {s->vel[s->vcur], 16ull * N}};
  for (auto& f : fields) {
    void* p = nullptr;
    if (hipMalloc(&p, f.second) != hipSuccess) {
      for (void* q : save) hipFree(q);
      return fail(s, XPBI_ERR_NOMEM, "autotune: device allocation failed");
    }
    hipMemcpyAsync(p, f.first, f.second, hipMemcpyDeviceToDevice, s->stream);
    save.push_back(p);
  }
  const int cur0 = s->cur, vcur0 = s->vcur;
  const int oldBlocks = s->persist ? (int)s->grid : 0;
  std::vector<int> cands = {0, cus, 2 * cus, 4 * cus};
  int best = 0;
  float bestMs = 1e30f;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  for (int c : cands) {
    if (c > 0 && (u32)c >= s->nb) continue;
    xpbi_set_launch(h, c);
    // restore, warm one substep (graph capture for this shape), then time
    s->cur = cur0;
    s->vcur = vcur0;
    for (size_t k = 0; k < save.size(); k++) hipMemcpyAsync(fields[k].first, save[k], fields[k].second,
                                                           hipMemcpyDeviceToDevice, s->stream);
    if ((r = xpbi_graph_step(h, 2, iterations, has_water))) break;
    hipEventRecord(e0, s->stream);
    if ((r = xpbi_graph_step(h, 2 * substeps, iterations, has_water))) break;
    hipEventRecord(e1, s->stream);
    hipEventSynchronize(e1);
    float ms = 0.0f;
    hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)(2 * substeps);
    if (ms < bestMs) {
      bestMs = ms;
      best = c;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  // the simulation continues from where it was, with the fastest shape
  xpbi_sync(h);
  s->cur = cur0;
  s->vcur = vcur0;
  for (size_t k = 0; k < save.size(); k++)
    hipMemcpy(fields[k].first, save[k], fields[k].second, hipMemcpyDeviceToDevice);
  for (void* q : save) hipFree(q);
  drop_graph(s);
  xpbi_set_launch(h, r ? oldBlocks : best);
  if (r) return r;
  if (best_blocks) *best_blocks = best;
  if (best_ms) *best_ms = bestMs;
  return XPBI_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free substep
// xpbi_substep_noalloc: one substep of a single scene on the caller's buffers,
// every intermediate in `scratch`, a static device array of this library (each
// device holds its own copy until the process ends).  Host code never uses the
// array's name as a device address: its address on each device comes from
// hipGetSymbolAddress, taken once per device with that device current.
namespace {

constexpr u32 kNaMaxN = XPBI_NOALLOC_MAX_PARTICLES;
constexpr u32 kNaMaxC = XPBI_NOALLOC_MAX_CELLS;
constexpr u32 kNaSlots = 64;  // support-neighbour list slots (more neighbours: the row walk)
constexpr size_t kNaAlign = 256;
constexpr size_t naUp(size_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }
// the scratch layout: per-particle arrays for kNaMaxN particles, per-cell arrays
// for kNaMaxC cells, then the scan partials and the scene's tables
constexpr size_t kNaPerParticle = 2 * (16 + 16 + 16 + 4 + 16 + 4)  // the two particle sets
                                  + 2 * 16                          // velocities
                                  + 2 * 4 + 2 * 4                   // sort keys and values
                                  + 16 + 16 + 4                     // L
                                  + 16 + 16 + 16 + 16               // Sp, selfDv
                                  + 2 * kNaSlots + 8                // neighbour list and heads
                                  + 16;                             // predicted positions
// per-particle and per-cell arrays at the maximum sizes, plus 1 MiB for the
// scan partials, the tables and the alignment of each array
constexpr size_t kNaScratchBytes = kNaPerParticle * (size_t)kNaMaxN + 8ull * kNaMaxC + (1ull << 20);

__device__ __align__(256) unsigned char scratch[kNaScratchBytes];

// the scratch address on each device (set once per device, under naLock)
constexpr int kNaMaxDevices = 64;
unsigned char* g_naPtr[kNaMaxDevices] = {};
std::mutex g_naLock;

struct NaTables {
  KP kp;
  DevMat mats[8];
};

// the scene's tables into scratch: its KP record, 8 materials and the work table
// (one entry per run of 256 particles)
__global__ __launch_bounds__(256) void k_na_tables(NaTables t, KP* __restrict__ kp, DevMat* __restrict__ mats,
                                                   uint2* __restrict__ tab, u32 nb) {
  if (threadIdx.x == 0) *kp = t.kp;
  if (threadIdx.x < 8) mats[threadIdx.x] = t.mats[threadIdx.x];
  for (u32 e = threadIdx.x; e < nb; e += 256u) tab[e] = make_uint2(0u, 256u * e);
}

// caller layouts (pos, vel, aux: float4; F: three vec4 columns) -> set 0 and
// velocity 0; ids 0..N-1
__global__ __launch_bounds__(256) void k_na_load(const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                 const float4* __restrict__ F, const float4* __restrict__ aux, u32 N,
                                                 PSet d, float4* __restrict__ vd) {
  const u32 i = blockIdx.x * 256u + threadIdx.x;
  if (i >= N) return;
  d.pos[i] = pos[i];
  vd[i] = vel[i];
  d.aux[i] = aux[i];
  const float4 c0 = F[3 * (size_t)i], c1 = F[3 * (size_t)i + 1], c2 = F[3 * (size_t)i + 2];
  d.Fa[i] = make_float4(c0.x, c0.y, c0.z, c1.x);
  d.Fb[i] = make_float4(c1.y, c1.z, c2.x, c2.y);
  d.Fc[i] = c2.z;
  d.id[i] = i;
}

// the substep's final set and velocity -> the caller's buffers (pad lanes of F 0)
__global__ __launch_bounds__(256) void k_na_store(PSet s, const float4* __restrict__ vs, u32 N,
                                                  float4* __restrict__ pos, float4* __restrict__ vel,
                                                  float4* __restrict__ F, float4* __restrict__ aux) {
  const u32 i = blockIdx.x * 256u + threadIdx.x;
  if (i >= N) return;
  pos[i] = s.pos[i];
  vel[i] = vs[i];
  aux[i] = s.aux[i];
  const float4 a = s.Fa[i], b = s.Fb[i];
  F[3 * (size_t)i] = make_float4(a.x, a.y, a.z, 0.0f);
  F[3 * (size_t)i + 1] = make_float4(a.w, b.x, b.y, 0.0f);
  F[3 * (size_t)i + 2] = make_float4(b.z, b.w, s.Fc[i], 0.0f);
}

// the scratch address on the current device (taken on the first call there)
int na_scratch(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kNaMaxDevices) return fail(nullptr, XPBI_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> g(g_naLock);
  if (!g_naPtr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(nullptr, e, "hipGetSymbolAddress(scratch)");
    g_naPtr[dev] = (unsigned char*)p;
  }
  *out = g_naPtr[dev];
  return XPBI_OK;
}

}  // namespace

extern "C" {

size_t xpbi_noalloc_scratch_bytes(void) { return kNaScratchBytes; }

int xpbi_substep_noalloc(const xpbi_params* p, const xpbi_material* mats, int n, void* pos, void* vel, void* F,
                         void* aux, int iterations, int has_water, void* stream) {
  if (!p || (n > 0 && !mats)) return fail(nullptr, XPBI_ERR_ARG, "params or materials is null");
  if (n < 0 || n > 8) return fail(nullptr, XPBI_ERR_ARG, "material count must be 0..8");
  const u32 N = p->N, C = p->numCells;
  if (N > kNaMaxN) return fail(nullptr, XPBI_ERR_ARG, "more particles than the allocation-free substep holds");
  if (C > kNaMaxC) return fail(nullptr, XPBI_ERR_ARG, "more cells than the allocation-free substep holds");
  if (iterations < 0) return fail(nullptr, XPBI_ERR_ARG, "iterations < 0");
  if (p->gridDims[0] < 1 || p->gridDims[1] < 1 || p->gridDims[2] < 1 ||
      (u64)p->gridDims[0] * (u64)p->gridDims[1] * (u64)p->gridDims[2] != (u64)C)
    return fail(nullptr, XPBI_ERR_ARG, "gridDims do not multiply to numCells");
  if (!(p->dt > 0.0f) || !(p->h > 0.0f)) return fail(nullptr, XPBI_ERR_ARG, "dt and h must be > 0");
  if (!(p->support > 0.0f) || p->support > p->cellSize)
    return fail(nullptr, XPBI_ERR_ARG, "support must be in (0, cellSize]");
  for (void* b : {pos, vel, F, aux})
    if (N > 0 && (!b || ((uintptr_t)b & 15u))) return fail(nullptr, XPBI_ERR_ARG, "state buffers must be 16-byte aligned");
  if (N == 0) return XPBI_OK;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDevice");
  unsigned char* base = nullptr;
  int r = na_scratch(dev, &base);
  if (r) return r;

  // carve the scratch (the same arrays xpbi_create allocates, for this N and C)
  size_t off = 0;
  auto take = [&](size_t bytes) {
    void* q = base + off;
    off += naUp(bytes);
    return q;
  };
  Sim s;  // host bookkeeping only: every pointer below is into scratch
  s.dev = dev;
  s.stream = (hipStream_t)stream;
  s.N = N;
  s.C = C;
  s.nb = (N + 255) / 256;
  s.grid = s.nb;
  s.persist = false;
  for (int k = 0; k < 2; k++) {
    s.set[k].pos = (float4*)take(16ull * N);
    s.set[k].Fa = (float4*)take(16ull * N);
    s.set[k].Fb = (float4*)take(16ull * N);
    s.set[k].Fc = (float*)take(4ull * N);
    s.set[k].aux = (float4*)take(16ull * N);
    s.set[k].id = (u32*)take(4ull * N);
    s.vel[k] = (float4*)take(16ull * N);
    s.keys[k] = (u32*)take(4ull * N);
    s.vals[k] = (u32*)take(4ull * N);
  }
  s.L = M3P{(float4*)take(16ull * N), (float4*)take(16ull * N), (float*)take(4ull * N)};
  s.Sp = SpP{(float4*)take(16ull * N), (float4*)take(16ull * N), (float4*)take(16ull * N)};
  s.selfDv = (float4*)take(16ull * N);
  s.nbrK = kNaSlots;
  s.nbrLst = (uint16_t*)take(2ull * kNaSlots * N);
  s.nbrHead = (uint2*)take(8ull * N);
  s.xc = (float4*)take(16ull * N);
  s.cellStart = (u32*)take(4ull * ((size_t)C + 1));
  s.count = (u32*)take(4ull * C);
  s.scanPart = (u32*)take(4ull * scan_partials_needed(C));
  s.dKP = (KP*)take(sizeof(KP));
  s.vmaxSlot = (u32*)take(2 * sizeof(u32));
  s.dMats = (DevMat*)take(8 * sizeof(DevMat));
  s.dTab = (uint2*)take(sizeof(uint2) * s.nb);
  if (off > kNaScratchBytes) return fail(nullptr, XPBI_ERR_ARG, "scratch layout exceeds the scratch array");
  s.perm = s.vals[0];
  s.sortedKeys = s.keys[0];

  Scene sc;
  sc.N = N;
  sc.C = C;
  sc.params = *p;
  build_kp(sc);
  NaTables t;
  t.kp = sc.kp;
  std::memset(t.mats, 0, sizeof(t.mats));
  for (int k = 0; k < n; k++) std::memcpy(&t.mats[k], &mats[k], sizeof(DevMat));
  const u32 g = (N + 255) / 256;
  hipLaunchKernelGGL(k_na_tables, dim3(1), dim3(256), 0, s.stream, t, s.dKP, s.dMats, s.dTab, s.nb);
  hipLaunchKernelGGL(k_na_load, dim3(g), dim3(256), 0, s.stream, (const float4*)pos, (const float4*)vel,
                     (const float4*)F, (const float4*)aux, N, s.set[0], s.vel[0]);
  launch_sort(&s);
  launch_reorder(&s);
  launch_solve_stages(&s, iterations, has_water != 0, XPBI_STAGE_FULL);
  hipLaunchKernelGGL(k_na_store, dim3(g), dim3(256), 0, s.stream, s.set[s.cur], (const float4*)s.vel[s.vcur], N,
                     (float4*)pos, (float4*)vel, (float4*)F, (float4*)aux);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(nullptr, e, "xpbi_substep_noalloc");
  return XPBI_OK;
}

}  // extern "C"
