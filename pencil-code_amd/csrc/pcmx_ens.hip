// pcmx_ens.hip -- ensembles of independent periodic boxes (C ABI pcmx_ens_*,
// declared in include/pcmx.h).
//
// All boxes share the physics (variables, EOS, precision) and each has its
// own grid, spacing, viscosity, resistivity and fixed time step.  Their
// states live in one allocation per array ([box][var][z][y][x], interior
// only), and a substep of the whole ensemble is one persistent launch of the
// split z-marching kernel over a work list of (box, tile, z-chunk) items,
// items of a box contiguous.  Two RK3 steps (six launches, after which the
// ping-pong buffers are back where they started) are captured once in a HIP
// graph and replayed.  pcmx_ens_tune picks the tile shape, the chunk length
// and the blocks per CU by timing the candidates on the loaded state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pcmx_hostutil.h"
#include "pcmx_launch.h"
#include "pcmx_zsplit.h"

namespace pcmx {

// Block b reduces the partials of box b's work items [wbeg[b], wbeg[b+1]) in
// a fixed order: urms, umax, brms (diagnostics.f90 sum_mn_name / max_mn_name
// with lsqrt over the box) and the box means of every variable.
__global__ void __launch_bounds__(256) k_ens_finalize(const double* partials, const int* wbeg, double* results,
                                                      const double* ncells, int mvar) {
  __shared__ double lds[4][NRED];
  const int box = blockIdx.x;
  double r[NRED];
#pragma unroll
  for (int q = 0; q < NRED; ++q) r[q] = 0.0;
  for (long b = wbeg[box] + threadIdx.x; b < wbeg[box + 1]; b += 256) {
#pragma unroll
    for (int q = 0; q < NRED; ++q) {
      const double v = partials[b * NRED + q];
      r[q] = (q == 1 || q == 3) ? fmax(r[q], v) : r[q] + v;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NRED; ++q) {
    const double v = (q == 1 || q == 3) ? wave_max(r[q]) : wave_sum(r[q]);
    if (lane == 0) lds[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[NRED];
    for (int q = 0; q < NRED; ++q) {
      t[q] = lds[0][q];
      for (int wv = 1; wv < 4; ++wv) t[q] = (q == 1 || q == 3) ? fmax(t[q], lds[wv][q]) : t[q] + lds[wv][q];
    }
    double* res = results + (long)box * R_COUNT;
    res[R_URMS] = sqrt(t[0] / ncells[box]);
    res[R_UMAX] = sqrt(t[1]);
    res[R_BRMS] = sqrt(t[2] / ncells[box]);
    for (int v = 0; v < mvar; ++v) res[R_MEAN0 + v] = t[4 + v] / ncells[box];
  }
}

}  // namespace pcmx

using namespace pcmx;

namespace {

struct Ens {
  pcmx_params par{};  // shared physics; grid fields are those of box 0
  int nbox = 0, nvar = 0;
  bool mag = false, ent = false, mixed = false;
  std::vector<pcmx_box> box;
  std::vector<long> off, ncell;  // element offset of each box (all variables), cells per box
  long total = 0;                // elements of one state array
  double* f[2] = {nullptr, nullptr};
  double* w = nullptr;
  double* partials = nullptr;
  long partials_cap = 0;
  double* results = nullptr;     // nbox * R_COUNT
  double* d_ncell = nullptr;
  int* d_wbeg = nullptr;
  EnsBox* d_cfg = nullptr;       // [parity 2][substep 3] + DT stage: 7 * nbox
  EnsWork* d_work = nullptr;
  long work_cap = 0;
  int nwork = 0;
  int tile = TILE_32x8, lzmax = 32, bpc = 1, nblocks = 0;
  int cur = 0;
  hipStream_t stream = nullptr;
  bool use_graph = true;
  hipGraphExec_t exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [ldiag][cur]
  int ncu = 256;
  bool timing = false;
  double timing_ms = 0.0;
  long timing_n = 0;
};

Ens* ge = nullptr;

int require_ens() { return ge ? PCMX_OK : fail(PCMX_ERR_STATE, "pcmx_ens_create has not been called"); }

void drop_graphs(Ens& e) {
  for (auto& row : e.exec)
    for (auto& x : row)
      if (x) {
        (void)hipGraphExecDestroy(x);
        x = nullptr;
      }
}

void free_ens(Ens* e) {
  if (!e) return;
  drop_graphs(*e);
  for (double* p : {e->f[0], e->f[1], e->w, e->partials, e->results, e->d_ncell})
    if (p) (void)hipFree(p);
  if (e->d_wbeg) (void)hipFree(e->d_wbeg);
  if (e->d_cfg) (void)hipFree(e->d_cfg);
  if (e->d_work) (void)hipFree(e->d_work);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

pcmx_params box_params(const Ens& e, int b) {
  pcmx_params p = e.par;
  p.nx = e.box[b].nx;
  p.ny = e.box[b].ny;
  p.nz = e.box[b].nz;
  p.dx = e.box[b].dx;
  p.dy = e.box[b].dy;
  p.dz = e.box[b].dz;
  p.nu = e.box[b].nu;
  p.eta = e.box[b].eta;
  return p;
}

bool tile_fits(const Ens& e, int tile) {
  for (const auto& b : e.box)
    if (b.nx % tile_tx(tile) || b.ny % tile_ty(tile) || (long)b.nx * b.ny >= (1L << 31)) return false;
  return true;
}

// chunk length of a box: the largest divisor of nz not above lzmax (at
// least 8 planes unless nz is smaller)
int box_lz(int nz, int lzmax) {
  for (int d = std::min(nz, lzmax); d >= 1; --d)
    if (nz % d == 0 && (d >= 8 || d == nz)) return d;
  return nz;
}

long count_work(const Ens& e, int tile, int lzmax) {
  long n = 0;
  for (const auto& b : e.box) n += (long)(b.nx / tile_tx(tile)) * (b.ny / tile_ty(tile)) * (b.nz / box_lz(b.nz, lzmax));
  return n;
}

// Builds the work list and the per-(parity, substep) launch arguments for a
// tile shape, chunk length and blocks per CU, and uploads them.
int build_plan(Ens& e, int tile, int lzmax, int bpc) {
  drop_graphs(e);
  const int tx = tile_tx(tile), ty = tile_ty(tile);
  std::vector<EnsWork> work;
  std::vector<int> wbeg(e.nbox + 1, 0);
  for (int b = 0; b < e.nbox; ++b) {
    const pcmx_box& B = e.box[b];
    const int lz = box_lz(B.nz, lzmax);
    wbeg[b] = (int)work.size();
    for (int z0 = 0; z0 < B.nz; z0 += lz)
      for (int y0 = 0; y0 < B.ny; y0 += ty)
        for (int x0 = 0; x0 < B.nx; x0 += tx) work.push_back(EnsWork{b, x0, y0, z0, lz});
  }
  wbeg[e.nbox] = (int)work.size();
  if ((long)work.size() > e.work_cap)
    return fail(PCMX_ERR_ARG, "work list of %zu items exceeds the %ld allocated", work.size(), e.work_cap);
  e.nwork = (int)work.size();
  e.tile = tile;
  e.lzmax = lzmax;
  e.bpc = bpc;
  e.nblocks = std::max(1, std::min(e.nwork, e.ncu * bpc));
  // launch arguments: [parity][substep] for the explicit 2N-RK3 step, then
  // the reductions-only stage on f[cur]
  std::vector<EnsBox> cfg(7L * e.nbox);
  for (int par = 0; par < 2; ++par)
    for (int s = 0; s < 3; ++s)
      for (int b = 0; b < e.nbox; ++b) {
        // substep s of a step starting in buffer `par`: in = par ^ (s & 1)
        const int in = par ^ (s & 1);
        EnsBox& c = cfg[(par * 3 + s) * (long)e.nbox + b];
        KArgs& a = c.a;
        std::memset(&a, 0, sizeof(a));
        a.fin = e.f[in] + e.off[b];
        a.fout = e.f[1 - in] + e.off[b];
        a.w = e.w + e.off[b];
        a.wout = a.w;
        a.means = e.results + (long)b * R_COUNT + R_MEAN0;
        a.nx = e.box[b].nx;
        a.ny = e.box[b].ny;
        a.nz = e.box[b].nz;
        a.mvar = e.nvar;
        a.iss = e.par.iss;
        a.iax = e.par.iax;
        a.nxyz = e.ncell[b];
        a.alpha = ALPHA_TS[s];
        a.beta = BETA_TS[s];
        a.dtb = e.box[b].dt * BETA_TS[s];  // dt_beta_ts (timestep.f90:150,215)
        a.first = (s == 0);
        a.write_w = (s < 2);
        a.mode = MODE_UPDATE;
        a.stage = (s == 0) ? ST_S1 : (s == 1 ? ST_S2 : ST_S3);
        c.P = make_phys(box_params(e, b));
      }
  for (int b = 0; b < e.nbox; ++b) cfg[6L * e.nbox + b] = cfg[(long)b];  // DT stage: f[0], set per call
  HIPCHECK(hipMemcpy(e.d_work, work.data(), sizeof(EnsWork) * work.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(e.d_wbeg, wbeg.data(), sizeof(int) * wbeg.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(e.d_cfg, cfg.data(), sizeof(EnsBox) * cfg.size(), hipMemcpyHostToDevice));
  return PCMX_OK;
}

LaunchCfg ens_cfg(const Ens& e) {
  LaunchCfg c{};
  c.kernel = 3;
  c.tile = e.tile;
  c.mag = e.mag;
  c.ent = e.ent;
  return c;
}

int launch_stage(Ens& e, bool mixed, int stage, const EnsBox* cfg, double* partials) {
  const LaunchCfg c = ens_cfg(e);
  const bool ok = mixed ? launch_ens_f32(c, stage, cfg, e.d_work, e.nwork, partials, e.nblocks, e.stream)
                        : launch_ens_f64(c, stage, cfg, e.d_work, e.nwork, partials, e.nblocks, e.stream);
  if (!ok) return fail(PCMX_ERR_ARG, "no ensemble kernel for this physics (isothermal hydro or MHD only)");
  HIPCHECK(hipGetLastError());
  return PCMX_OK;
}

int finalize(Ens& e) {
  hipLaunchKernelGGL(k_ens_finalize, dim3(e.nbox), dim3(256), 0, e.stream, e.partials, e.d_wbeg, e.results,
                     e.d_ncell, e.nvar);
  HIPCHECK(hipGetLastError());
  return PCMX_OK;
}

// One RK3 step of every box starting in buffer e.cur; enqueued on e.stream.
int enqueue_step(Ens& e, bool ldiag) {
  for (int s = 0; s < 3; ++s) {
    const EnsBox* cfg = e.d_cfg + (long)(e.cur * 3 + s) * e.nbox;
    const bool red = (s == 0 && ldiag);
    int rc = launch_stage(e, e.mixed, s == 0 ? ST_S1 : (s == 1 ? ST_S2 : ST_S3), cfg, red ? e.partials : nullptr);
    if (rc) return rc;
    if (red && (rc = finalize(e))) return rc;
  }
  e.cur = 1 - e.cur;
  return PCMX_OK;
}

// Box means (for the mixed form's offsets) and diagnostics of the current
// state: the FP64 reductions-only stage over f[cur].
int refresh_means(Ens& e) {
  std::vector<EnsBox> dtc(e.nbox);
  HIPCHECK(hipMemcpy(dtc.data(), e.d_cfg + (long)(e.cur * 3) * e.nbox, sizeof(EnsBox) * e.nbox,
                     hipMemcpyDeviceToHost));
  for (auto& c : dtc) {
    c.a.stage = ST_DT;
    c.a.mode = MODE_DT;
  }
  EnsBox* d = e.d_cfg + 6L * e.nbox;
  HIPCHECK(hipMemcpy(d, dtc.data(), sizeof(EnsBox) * e.nbox, hipMemcpyHostToDevice));
  int rc = launch_stage(e, false, ST_DT, d, e.partials);
  if (rc) return rc;
  if ((rc = finalize(e))) return rc;
  HIPCHECK(hipStreamSynchronize(e.stream));
  return PCMX_OK;
}

int run_steps(Ens& e, int nsteps, bool ldiag) {
  int done = 0;
  if (e.use_graph) {
    while (nsteps - done >= 2) {
      hipGraphExec_t& x = e.exec[ldiag ? 1 : 0][e.cur];
      if (!x) {
        const int cur0 = e.cur;
        hipGraph_t graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_step(e, ldiag);
        if (!rc) rc = enqueue_step(e, ldiag);
        hipError_t ce = hipStreamEndCapture(e.stream, &graph);
        e.cur = cur0;
        if (rc) {
          if (graph) (void)hipGraphDestroy(graph);
          return rc;
        }
        if (ce != hipSuccess) return fail(PCMX_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ce));
        hipError_t ie = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return fail(PCMX_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
      }
      HIPCHECK(hipGraphLaunch(x, e.stream));
      done += 2;  // two steps bring the buffers back: e.cur unchanged
    }
  }
  for (; done < nsteps; ++done) {
    int rc = enqueue_step(e, ldiag);
    if (rc) return rc;
  }
  return PCMX_OK;
}

}  // namespace

extern "C" {

int pcmx_ens_create(const pcmx_params* par, const pcmx_box* boxes, int nbox, int flags) {
  if (!par || !boxes || nbox < 1) return fail(PCMX_ERR_ARG, "pcmx_ens_create needs params and at least one box");
  if (par->lupw_uu || par->lupw_lnrho) return fail(PCMX_ERR_ARG, "the ensemble does not provide upwinding");
  if (!par->lperi_z || par->gravz != 0.0)
    return fail(PCMX_ERR_ARG, "the ensemble runs periodic boxes without gravity");
  if (layers_on(*par))
    return fail(PCMX_ERR_ARG, "the ensemble does not provide rotation, K-profile conduction, heating or cooling");
  if (ge) {
    free_ens(ge);
    ge = nullptr;
  }
  Ens* e = new Ens();
  e->par = *par;
  e->nbox = nbox;
  e->box.assign(boxes, boxes + nbox);
  e->nvar = par->mvar;
  e->mag = par->iax >= 0;
  e->ent = par->iss >= 0;
  e->mixed = par->precision == PCMX_PREC_MIXED;
  e->use_graph = !(flags & PCMX_ENS_NO_GRAPH);
  int rc;
  for (int b = 0; b < nbox; ++b) {
    const pcmx_params p = box_params(*e, b);
    if ((rc = check_params(&p))) {
      delete e;
      return fail(rc, "box %d: %s", b, last_error());
    }
    if (!(boxes[b].dt > 0)) {
      delete e;
      return fail(PCMX_ERR_ARG, "box %d: the ensemble needs a positive fixed dt", b);
    }
  }
  if (par->ldt || e->ent) {
    delete e;
    return fail(PCMX_ERR_ARG, "the ensemble runs isothermal hydro or MHD with fixed steps");
  }
  if (!tile_fits(*e, TILE_32x8) && !tile_fits(*e, TILE_16x16)) {
    delete e;
    return fail(PCMX_ERR_ARG, "every box needs nx, ny multiples of 16 (or nx of 32 and ny of 8)");
  }
  e->off.resize(nbox);
  e->ncell.resize(nbox);
  std::vector<double> nc(nbox);
  for (int b = 0; b < nbox; ++b) {
    e->ncell[b] = (long)boxes[b].nx * boxes[b].ny * boxes[b].nz;
    e->off[b] = e->total;
    e->total += e->ncell[b] * e->nvar;
    nc[b] = (double)e->ncell[b];
  }
  // work list capacity: the smallest tiles and chunks any plan uses
  e->work_cap = std::max(count_work(*e, TILE_16x16, 8), count_work(*e, TILE_32x8, 8));
  e->partials_cap = e->work_cap;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&e->ncu, hipDeviceAttributeMultiprocessorCount, dev);
  auto alloc = [&](void** p, size_t n) -> int {
    hipError_t err = hipMalloc(p, n);
    return err == hipSuccess ? PCMX_OK : fail(PCMX_ERR_HIP, "hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(err));
  };
  const size_t bytes = sizeof(double) * e->total;
  if ((rc = alloc((void**)&e->f[0], bytes)) || (rc = alloc((void**)&e->f[1], bytes)) ||
      (rc = alloc((void**)&e->w, bytes)) || (rc = alloc((void**)&e->partials, sizeof(double) * NRED * e->partials_cap)) ||
      (rc = alloc((void**)&e->results, sizeof(double) * R_COUNT * nbox)) ||
      (rc = alloc((void**)&e->d_ncell, sizeof(double) * nbox)) ||
      (rc = alloc((void**)&e->d_wbeg, sizeof(int) * (nbox + 1))) ||
      (rc = alloc((void**)&e->d_cfg, sizeof(EnsBox) * 7 * nbox)) ||
      (rc = alloc((void**)&e->d_work, sizeof(EnsWork) * e->work_cap))) {
    free_ens(e);
    return rc;
  }
  hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (err != hipSuccess) {
    free_ens(e);
    return fail(PCMX_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(err));
  }
  (void)hipMemsetAsync(e->w, 0, bytes, e->stream);
  (void)hipMemsetAsync(e->f[0], 0, bytes, e->stream);
  (void)hipMemsetAsync(e->results, 0, sizeof(double) * R_COUNT * nbox, e->stream);
  if (hipMemcpyAsync(e->d_ncell, nc.data(), sizeof(double) * nbox, hipMemcpyHostToDevice, e->stream) != hipSuccess ||
      hipStreamSynchronize(e->stream) != hipSuccess) {
    free_ens(e);
    return fail(PCMX_ERR_HIP, "ensemble initialisation failed");
  }
  const int tile = tile_fits(*e, TILE_32x8) ? TILE_32x8 : TILE_16x16;
  if ((rc = build_plan(*e, tile, 32, 1))) {
    free_ens(e);
    return rc;
  }
  ge = e;
  return PCMX_OK;
}

int pcmx_ens_destroy(void) {
  if (ge) {
    (void)hipStreamSynchronize(ge->stream);
    free_ens(ge);
    ge = nullptr;
  }
  return PCMX_OK;
}

int pcmx_ens_load(int ibox, const double* f) {
  int rc = require_ens();
  if (rc) return rc;
  Ens& e = *ge;
  if (ibox < 0 || ibox >= e.nbox || !f) return fail(PCMX_ERR_ARG, "pcmx_ens_load: bad box %d or NULL state", ibox);
  HIPCHECK(hipMemcpyAsync(e.f[e.cur] + e.off[ibox], f, sizeof(double) * e.ncell[ibox] * e.nvar,
                          hipMemcpyHostToDevice, e.stream));
  HIPCHECK(hipStreamSynchronize(e.stream));
  return refresh_means(e);
}

int pcmx_ens_copy(int ibox, double* f) {
  int rc = require_ens();
  if (rc) return rc;
  Ens& e = *ge;
  if (ibox < 0 || ibox >= e.nbox || !f) return fail(PCMX_ERR_ARG, "pcmx_ens_copy: bad box %d or NULL state", ibox);
  HIPCHECK(hipMemcpyAsync(f, e.f[e.cur] + e.off[ibox], sizeof(double) * e.ncell[ibox] * e.nvar,
                          hipMemcpyDeviceToHost, e.stream));
  HIPCHECK(hipStreamSynchronize(e.stream));
  return PCMX_OK;
}

int pcmx_ens_step(int nsteps, int ldiag) {
  int rc = require_ens();
  if (rc) return rc;
  Ens& e = *ge;
  if (nsteps < 0) return fail(PCMX_ERR_ARG, "nsteps < 0");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e.timing) {
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, e.stream));
  }
  if ((rc = run_steps(e, nsteps, ldiag != 0))) return rc;
  if (e.timing) {
    HIPCHECK(hipEventRecord(e1, e.stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    e.timing_ms = ms;
    e.timing_n = 3L * nsteps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  HIPCHECK(hipStreamSynchronize(e.stream));
  return PCMX_OK;
}

int pcmx_ens_reduced(double* dst) {
  int rc = require_ens();
  if (rc) return rc;
  if (!dst) return fail(PCMX_ERR_ARG, "dst is NULL");
  Ens& e = *ge;
  std::vector<double> r((size_t)R_COUNT * e.nbox);
  HIPCHECK(hipMemcpy(r.data(), e.results, sizeof(double) * r.size(), hipMemcpyDeviceToHost));
  for (int b = 0; b < e.nbox; ++b) {
    dst[3 * b + 0] = r[(size_t)b * R_COUNT + R_URMS];
    dst[3 * b + 1] = r[(size_t)b * R_COUNT + R_UMAX];
    dst[3 * b + 2] = r[(size_t)b * R_COUNT + R_BRMS];
  }
  return PCMX_OK;
}

int pcmx_ens_set_timing(int enable) {
  int rc = require_ens();
  if (rc) return rc;
  ge->timing = enable != 0;
  return PCMX_OK;
}

int pcmx_ens_timing(double* out) {
  int rc = require_ens();
  if (rc) return rc;
  if (!out) return fail(PCMX_ERR_ARG, "out is NULL");
  out[0] = ge->timing_n ? ge->timing_ms / ge->timing_n : 0.0;
  out[1] = (double)ge->timing_n;
  out[2] = ge->tile == TILE_16x16 ? 16.0 : 32.0;
  out[3] = ge->lzmax;
  out[4] = ge->bpc;
  out[5] = ge->nwork;
  return PCMX_OK;
}

int pcmx_ens_tune(int nsteps_probe) {
  int rc = require_ens();
  if (rc) return rc;
  Ens& e = *ge;
  if (nsteps_probe < 1) nsteps_probe = 2;
  // candidates are timed on a copy of the state, which is restored after
  double *sf = nullptr, *sw = nullptr;
  const size_t bytes = sizeof(double) * e.total;
  HIPCHECK(hipMalloc((void**)&sf, bytes));
  if (hipMalloc((void**)&sw, bytes) != hipSuccess) {
    (void)hipFree(sf);
    return fail(PCMX_ERR_HIP, "hipMalloc failed in pcmx_ens_tune");
  }
  HIPCHECK(hipMemcpyAsync(sf, e.f[e.cur], bytes, hipMemcpyDeviceToDevice, e.stream));
  HIPCHECK(hipMemcpyAsync(sw, e.w, bytes, hipMemcpyDeviceToDevice, e.stream));
  const int cur0 = e.cur;
  struct Cand { int tile, lz, bpc; };
  std::vector<Cand> cands;
  for (int tile : {TILE_32x8, TILE_16x16}) {
    if (!tile_fits(e, tile)) continue;
    for (int lz : {16, 32, 48, 96})
      for (int bpc : {1, 2}) cands.push_back({tile, lz, bpc});
  }
  float best = 1e30f;
  Cand bestc = {e.tile, e.lzmax, e.bpc};
  hipEvent_t t0, t1;
  HIPCHECK(hipEventCreate(&t0));
  HIPCHECK(hipEventCreate(&t1));
  for (const Cand& c : cands) {
    if (count_work(e, c.tile, c.lz) > e.work_cap) continue;
    if ((rc = build_plan(e, c.tile, c.lz, c.bpc))) break;
    for (int it = 0; it < nsteps_probe; ++it)
      if ((rc = enqueue_step(e, false))) break;          // warm up
    if (rc) break;
    HIPCHECK(hipEventRecord(t0, e.stream));
    for (int it = 0; it < nsteps_probe; ++it)
      if ((rc = enqueue_step(e, false))) break;
    if (rc) break;
    HIPCHECK(hipEventRecord(t1, e.stream));
    HIPCHECK(hipEventSynchronize(t1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, t0, t1));
    if (ms < best) {
      best = ms;
      bestc = c;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  e.cur = cur0;
  HIPCHECK(hipMemcpyAsync(e.f[e.cur], sf, bytes, hipMemcpyDeviceToDevice, e.stream));
  HIPCHECK(hipMemcpyAsync(e.w, sw, bytes, hipMemcpyDeviceToDevice, e.stream));
  HIPCHECK(hipStreamSynchronize(e.stream));
  (void)hipFree(sf);
  (void)hipFree(sw);
  if (rc) return rc;
  return build_plan(e, bestc.tile, bestc.lz, bestc.bpc);
}

}  // extern "C"
