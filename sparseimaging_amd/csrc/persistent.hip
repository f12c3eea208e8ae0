// persistent.hip -- host side of the persistent NUFFT MFISTA solver: launch
// configuration, workspace set-up, cooperative launches of single solves,
// batches and coupled movies.  The device code is in persistent_body.h (the
// gridding loop) and toeplitz_body.h (the Toeplitz loop of single solves).
#include "toeplitz_body.h"

namespace mf {

// ---------------------------------------------------------------------------
// Host side.

// Launch configuration shared by the problems of one launch.
struct PLaunch {
  const void* fn = nullptr;
  size_t sh = 0;       // dynamic LDS bytes
  int lds_tables = 0;  // tables staged in LDS
  int nblk = 0;        // workgroups of the launch
  int pair = 0;        // paired trial rounds for the L1+TSV problems of the launch
};

static int persistent_config(Context& C, int MSP, const std::vector<const NufftState*>& sets, bool batch,
                             bool allow_pair, PLaunch* L) {
  if ((MSP != 6 && MSP != 12) || sets.empty()) return MFISTA_E_ARG;
  // every problem of a launch uses the same gridding tile (one kernel)
  const int tx = sets[0]->geom.TX, ty = sets[0]->geom.TY;
  for (const NufftState* N : sets)
    if (N->geom.TX != tx || N->geom.TY != ty) return MFISTA_E_ARG;
  if (!tile_supported(tx, ty)) return MFISTA_E_ARG;
  // dynamic LDS: [twiddle + E4 tables, when they fit][degrid visibility
  // records][stage scratch: FP64 row/column FFT (two sequences per transform
  // in a paired round) or the batched gridder's entries].  Paired rounds are
  // used when their larger scratch still leaves the wanted number of
  // workgroups per CU.
  int dev = 0, ncu = 0, lds_max = 0;
  MF_CHECK(hipGetDevice(&dev));
  MF_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  MF_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t static_lds = 16 * kPW * 8 + sizeof(BlockState) + (kTD + 1) * 8 + sizeof(BarCtx) + 64;
  const int want = C.persist_blocks_per_cu > 0 ? C.persist_blocks_per_cu : 2;
  for (int pair = (C.persist_pair && allow_pair) ? 1 : 0; pair >= 0; pair--) {
    size_t sh = 0, tables = 0;
    for (const NufftState* N : sets) {
      const GridGeom& g = N->geom;
      const int Lm = g.Lx > g.Ly ? g.Lx : g.Ly;
      size_t s1 = (size_t)(pair ? 4 : 2) * Lm * sizeof(double2);
      if (grid_lds_bytes(g.TX, g.TY) > s1) s1 = grid_lds_bytes(g.TX, g.TY);
      s1 += kDVisCap * sizeof(DVis);
      if (s1 > sh) sh = s1;
      const size_t t1 = (table_lds_bytes(g.Lx, g.Ly, g.NX, g.NY) + 255) & ~(size_t)255;
      if (t1 > tables) tables = t1;
    }
    L->lds_tables = (tables + sh + static_lds <= (size_t)lds_max) ? 1 : 0;
    if (L->lds_tables) sh += tables;
    L->sh = sh;
    const bool tab = L->lds_tables != 0;
    const int cpt = (tx * ty) / MF_BLOCK;
    L->fn = (cpt == 1) ? persistent_kernel<1>(MSP, tab, batch)
            : (cpt == 2) ? persistent_kernel<2>(MSP, tab, batch)
                         : persistent_kernel<4>(MSP, tab, batch);
    int per_cu = 0;
    MF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, L->fn, MF_BLOCK, sh));
    if (per_cu < 1 && !pair) return MFISTA_E_HIP;
    if (pair && per_cu < want) continue;
    L->pair = pair;
    L->nblk = ncu * (per_cu < want ? per_cu : want);
    return 0;
  }
  return MFISTA_E_HIP;
}

// Allocates and initialises one problem's workspace and fills its PArgs for
// nblk workgroups.  P.xinit and P.xout are device pointers.
static int persistent_prepare(NufftState& N, PersistWork& W, const RunParams& P, int nblk, int lds_tables,
                              int hier, int pair_ok, hipStream_t st, PArgs* out) {
  const GridGeom& g = N.geom;
  const int NX = P.NX, NY = P.NY, n = NX * NY;
  const int n_row = (NX + 1) / 2, n_ky = NY + 1, n_elem = elem_blocks(n);
  const int n_deg = degrid_blocks(N.M, N.MSP);
  const size_t rb = (size_t)n * 4;
  const size_t nky = (size_t)NY + 1;
  const int pair = (pair_ok && !P.tv_mode) ? 1 : 0;
  const size_t t64 = nky * NX * sizeof(double2), gf64 = ((size_t)NY + N.MSP) * g.Lx * sizeof(double2);
  if (W.px.ensure(2 * rb) || W.pz.ensure(2 * rb) || W.pxn.ensure(4 * rb) || W.g.ensure(rb) ||
      W.vmod.ensure(4 * (size_t)N.M * sizeof(double2)) || W.T64.ensure((pair ? 2 : 1) * t64) ||
      W.Gf64.ensure((pair ? 2 : 1) * gf64) ||
      W.Hg.ensure(nky * g.Lx * sizeof(float2)) || W.U.ensure((size_t)NX * nky * sizeof(float2)) ||
      W.part.ensure((size_t)nblk * kPW * 8) || W.state.ensure(sizeof(DevState)) ||
      W.cost.ensure((size_t)(P.maxiter > 0 ? P.maxiter : 1) * 8) || W.bar.ensure(kBarBytes) ||
      W.fgpc.ensure((size_t)2 * (P.fgp_iter > 0 ? P.fgp_iter : 1) * 8) ||
      W.granges.ensure((size_t)(nblk + 1) * 4))
    return MFISTA_E_HIP;
  if (P.tv_mode && W.fgp.ensure(7 * rb)) return MFISTA_E_HIP;
  {
    std::vector<double> cc(2 * (size_t)(P.fgp_iter > 0 ? P.fgp_iter : 1), 0.0);
    double t = 1;
    for (int k = 0; k < P.fgp_iter; k++) {
      const double tn = (1 + sqrt(1 + 4 * t * t)) / 2;
      cc[2 * k] = 1 + (t - 1) / tn;
      cc[2 * k + 1] = -(t - 1) / tn;
      t = tn;
    }
    MF_CHECK(hipMemcpyAsync(W.fgpc.p, cc.data(), cc.size() * 8, hipMemcpyHostToDevice, st));
  }
  // Gridding ranges: workgroup b takes entries [tot b / nblk, tot (b+1) / nblk)
  // of the tile-ordered entry sequence.  The tile chunk list is re-cut at those
  // boundaries, so every workgroup gathers the same number of entries (to one);
  // a cut piece keeps the tile's "first" flag only on its first piece and the
  // "last" flag only on its last, so a tile split between workgroups is
  // flushed with atomics.
  //
  // Deterministic mode (P.deterministic): ranges are cut at tile boundaries
  // only, so every tile is gathered by one workgroup, in entry order, and
  // stored without atomics.  The adjoint grid is then the same bit for bit on
  // every run and every launch size; the cost is a less even split of the
  // gridding work (a dense tile is not shared).
  std::vector<int4> pc;
  if (P.deterministic) {
    const std::vector<int4>& ch = N.h_chunks;
    std::vector<int> rg(nblk + 1, 0);
    long tot = 0;
    for (const int4& c4 : ch) tot += c4.z - c4.y;
    pc = ch;
    long G = 0;  // entries before chunk j
    int k = 1;
    for (int j = 0; j < (int)ch.size(); j++) {
      if (ch[j].w & 1)  // a tile starts here: workgroups whose share begins by now start at it
        while (k < nblk && tot * k / nblk <= G) rg[k++] = j;
      G += ch[j].z - ch[j].y;
    }
    while (k <= nblk) rg[k++] = (int)ch.size();
    if (W.pchunks.ensure(pc.size() * sizeof(int4) + 16)) return MFISTA_E_HIP;
    if (!pc.empty())
      MF_CHECK(hipMemcpy(W.pchunks.p, pc.data(), pc.size() * sizeof(int4), hipMemcpyHostToDevice));
    MF_CHECK(hipMemcpy(W.granges.p, rg.data(), (size_t)(nblk + 1) * 4, hipMemcpyHostToDevice));
  } else {
    const std::vector<int4>& ch = N.h_chunks;
    std::vector<int> rg(nblk + 1, 0);
    long tot = 0;
    for (const int4& c4 : ch) tot += c4.z - c4.y;
    auto bnd = [&](int k) -> long { return tot * k / nblk; };
    pc.reserve(ch.size() + nblk);
    long G = 0;  // entries before the current chunk
    int k = 1;
    for (const int4& c4 : ch) {
      const int len = c4.z - c4.y;
      int off = 0;
      while (off < len) {
        while (k < nblk && bnd(k) <= G + off) rg[k++] = (int)pc.size();
        const long next = (k < nblk) ? bnd(k) : G + len;
        const int end = (int)std::min<long>(len, next - G);
        const int fl = ((off == 0) ? (c4.w & 1) : 0) | ((end == len) ? (c4.w & 2) : 0);
        pc.push_back(make_int4(c4.x, c4.y + off, c4.y + end, fl));
        off = end;
      }
      G += len;
    }
    while (k <= nblk) rg[k++] = (int)pc.size();
    if (W.pchunks.ensure(pc.size() * sizeof(int4) + 16)) return MFISTA_E_HIP;
    if (!pc.empty())
      MF_CHECK(hipMemcpy(W.pchunks.p, pc.data(), pc.size() * sizeof(int4), hipMemcpyHostToDevice));
    MF_CHECK(hipMemcpy(W.granges.p, rg.data(), (size_t)(nblk + 1) * 4, hipMemcpyHostToDevice));
  }
  if (P.xinit != W.px.p)
    MF_CHECK(hipMemcpyAsync(W.px.p, P.xinit, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  MF_CHECK(hipMemsetAsync(W.bar.p, 0, kBarBytes, st));
  MF_CHECK(hipMemsetAsync(W.Hg.p, 0, nky * g.Lx * sizeof(float2), st));  // the gridder adds into zeros
  PArgs& A = *out;
  memset(&A, 0, sizeof(A));
  A.g = g;
  A.Px = N.Px;
  A.Py = N.Py;
  A.pf = N.pf();
  A.pd = N.pd();
  A.twxf = (const float2*)N.twxf.p;
  A.twyf = (const float2*)N.twyf.p;
  A.twxd = (const double2*)N.twxd.p;
  A.twyd = (const double2*)N.twyd.p;
  A.e4xf = (const float*)N.e4xf.p;
  A.e4yf = (const float*)N.e4yf.p;
  A.e4xd = (const double*)N.e4xd.p;
  A.e4yd = (const double*)N.e4yd.p;
  A.Hg = (float2*)W.Hg.p;
  A.U = (float2*)W.U.p;
  A.T64 = (double2*)W.T64.p;
  A.Gf64 = (double2*)W.Gf64.p;
  A.T64b = pair ? (double2*)((char*)W.T64.p + t64) : nullptr;
  A.Gf64b = pair ? (double2*)((char*)W.Gf64.p + gf64) : nullptr;
  A.pair = pair;
  for (int q = 0; q < 4; q++) A.vm[q] = (double2*)W.vmod.p + (size_t)q * N.M;
  A.chunks = (const int4*)W.pchunks.p;
  A.nchunks = (int)pc.size();
  A.granges = (const int*)W.granges.p;
  A.ent = (const unsigned*)N.ent.p;
  A.x = (double*)W.px.p;
  A.z = (double*)W.pz.p;
  A.xn[0] = (double*)W.pxn.p;
  A.xn[1] = (double*)W.pxn.p + n;
  A.gr = (float*)W.g.p;
  A.box = P.box_flag ? P.box : nullptr;
  if (P.tv_mode) {
    float* fb = (float*)W.fgp.p;
    A.rA = fb;
    A.sA = fb + n;
    A.rB = fb + 2 * (size_t)n;
    A.sB = fb + 3 * (size_t)n;
    A.pp = fb + 4 * (size_t)n;
    A.qq = fb + 5 * (size_t)n;
  }
  A.fgp_c = (const double*)W.fgpc.p;
  A.fgp_iter = P.tv_mode ? P.fgp_iter : 0;
  A.tv_mode = P.tv_mode;
  A.maxiter = P.maxiter;
  A.nonneg = P.nonneg;
  A.lambda_l1 = P.lambda_l1;
  A.lambda_tsv = P.tv_mode ? 0.0 : P.lambda_tsv;
  A.lambda_tv = P.tv_mode ? P.lambda_tv : 0.0;
  A.eps_abs = P.eps_abs;
  A.rel_tol = P.rel_tol;
  A.eta_first = P.eta_first;
  A.cinit = P.cinit;
  A.adj_scale = 1.0 / (8.0 * (double)NX * NY);
  A.part = (double*)W.part.p;
  A.n_deg = n_deg;
  A.n_row = n_row;
  A.n_ky = n_ky;
  A.n_elem = n_elem;
  A.lds_tables = lds_tables;
  A.nblk = nblk;
  A.cost = (double*)W.cost.p;
  A.S = (DevState*)W.state.p;
  A.bar = (unsigned long long*)W.bar.p;
  A.hier_barrier = hier;
  // FGP block shape: tiles of fgp_tw x (256 / fgp_tw) pixels, 0 = along the flat image
  A.fgp_tw = (P.fgp_tw == 16 || P.fgp_tw == 32 || P.fgp_tw == 64) ? P.fgp_tw : 0;
  return 0;
}

// Reads one problem's outcome after the launch and copies x to P.xout.
static int persistent_collect(PersistWork& W, const RunParams& P, hipStream_t st, RunStats* out) {
  unsigned long long hb[8];
  MF_CHECK(hipMemcpy(hb, W.bar.p, sizeof(hb), hipMemcpyDeviceToHost));
  if (hb[1]) {
    fprintf(stderr,
            "[mfista_amd] persistent solver: grid barrier timed out (workgroup %llu waited on word %llu "
            "for %llu, saw %llu, after %llu polls)\n",
            hb[7] & 0xfffff, hb[4], hb[5], hb[6], hb[7] >> 20);
    return MFISTA_E_HIP;
  }
  DevState h;
  MF_CHECK(hipMemcpy(&h, W.state.p, sizeof(h), hipMemcpyDeviceToHost));
  out->ITER = h.ITER;
  out->c = h.c;
  out->max_trials = h.max_trials;
  out->total_trials = h.total_trials;
  out->status = h.status;
  out->steps = h.total_trials;
  out->rounds = h.rounds;
  MF_CHECK(hipMemcpyAsync(P.xout, W.px.p, (size_t)P.NX * P.NY * 8, hipMemcpyDeviceToDevice, st));
  return 0;
}

// Returns kCoopRefused (nothing ran) when the runtime refuses the cooperative
// launch, which it does in a process while another process holds the device's
// cooperative-launch resources.
static int timed_launch(const PLaunch& L, void** args, int nblk, hipStream_t st, float* ms) {
  if (!ms) {  // untimed (the allocation-free entry point creates no events)
    const hipError_t le = hipLaunchCooperativeKernel(L.fn, dim3(nblk), dim3(MF_BLOCK), args, (unsigned)L.sh, st);
    if (le != hipSuccess) {
      (void)hipGetLastError();
      fprintf(stderr, "[mfista_amd] cooperative launch refused (%s)\n", hipGetErrorName(le));
      return kCoopRefused;
    }
    MF_CHECK(hipStreamSynchronize(st));
    MF_CHECK(hipGetLastError());
    return 0;
  }
  hipEvent_t e0, e1;
  MF_CHECK(hipEventCreate(&e0));
  MF_CHECK(hipEventCreate(&e1));
  MF_CHECK(hipEventRecord(e0, st));
  const hipError_t le = hipLaunchCooperativeKernel(L.fn, dim3(nblk), dim3(MF_BLOCK), args, (unsigned)L.sh, st);
  if (le != hipSuccess) {
    (void)hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    fprintf(stderr, "[mfista_amd] cooperative launch refused (%s)\n", hipGetErrorName(le));
    return kCoopRefused;
  }
  MF_CHECK(hipEventRecord(e1, st));
  MF_CHECK(hipStreamSynchronize(st));
  MF_CHECK(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  MF_CHECK(hipGetLastError());
  return 0;
}

// Trace summary (MFISTA_PERSISTENT_TRACE=<file>): per stage tag, the slowest
// workgroup's time from leaving the previous barrier to arriving at this one,
// and the barrier's tail (last arrival to last release), appended to the file.
static void write_trace(const char* path, const std::vector<unsigned long long>& tr, const std::vector<int>& tag,
                        int nbar, int nrow, double clock_khz, const RunStats& rs, const RunParams& P, int M,
                        bool toep = false) {
  static const char* grid_names[] = {"?",        "init-row", "init-col", "init-degrid", "P1-grid", "P2-colinv",
                                     "FGP-iter", "TV-c2r",   "T1-rows",  "T2-cols",     "T3-degrid"};
  static const char* toep_names[] = {"?", "init-row", "init-conv", "init-c2r", "momentum", "?",
                                     "?", "?",        "T1-rows",   "T2-conv",  "T3-c2r"};
  const char* const* names = toep ? toep_names : grid_names;
  double st[11] = {0}, tl[11] = {0};
  int cnt[11] = {0};
  const double us = 1e3 / clock_khz;
  for (int e = 1; e < nbar; e++) {
    unsigned long long amax = 0, rmax = 0;
    long long wmax = 0;
    for (int r = 0; r < nrow; r++) {
      const unsigned long long ta = tr[((size_t)e * nrow + r) * 2], rl = tr[((size_t)e * nrow + r) * 2 + 1];
      const unsigned long long prev = tr[((size_t)(e - 1) * nrow + r) * 2 + 1];
      amax = std::max(amax, ta);
      rmax = std::max(rmax, rl);
      wmax = std::max(wmax, (long long)(ta - prev));
    }
    const int k = (tag[e] >= 0 && tag[e] <= 10) ? tag[e] : 0;
    st[k] += wmax * us;
    tl[k] += (double)(rmax - amax) * us;
    cnt[k]++;
  }
  FILE* f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "persistent solve (%s loop) %dx%d M=%d: ITER %d trials %d rounds %d kernel %.3f ms, %d barriers traced\n",
          toep ? "Toeplitz" : "gridding", P.NX, P.NY, M, rs.ITER, rs.total_trials, rs.rounds, rs.kernel_ms, nbar);
  fprintf(f, "%-12s %7s %14s %13s %10s\n", "stage", "count", "mean stage us", "mean tail us", "total ms");
  double tot = 0;
  for (int k = 1; k <= 10; k++) {
    if (!cnt[k]) continue;
    fprintf(f, "%-12s %7d %14.2f %13.2f %10.3f\n", names[k], cnt[k], st[k] / cnt[k], tl[k] / cnt[k],
            (st[k] + tl[k]) * 1e-3);
    tot += st[k] + tl[k];
  }
  fprintf(f, "sum over traced barriers %.3f ms\n\n", tot * 1e-3);
  fclose(f);
}

// copies the barrier trace of the last single-solve launch back and appends its summary to `path`
static int read_trace(Context& C, const char* path, int tcap, int nblk, const RunStats& rs, const RunParams& P,
                      bool toep) {
  const long nbar = tcap;
  std::vector<unsigned long long> tr((size_t)nbar * nblk * 2);
  std::vector<int> tag(nbar);
  MF_CHECK(hipMemcpy(tr.data(), C.trace.p, tr.size() * 8, hipMemcpyDeviceToHost));
  MF_CHECK(hipMemcpy(tag.data(), C.trace_tag.p, tag.size() * 4, hipMemcpyDeviceToHost));
  int dev = 0, khz = 0;
  MF_CHECK(hipGetDevice(&dev));
  MF_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  int used = 0;  // barriers actually passed: their release clocks are set
  while (used < nbar && tr[((size_t)used * nblk) * 2 + 1] != 0) used++;
  write_trace(path, tr, tag, used, nblk, khz > 0 ? khz : 100000, rs, P, C.nu.M, toep);
  return 0;
}

int solve_nufft_persistent(Context& C, const RunParams& P, RunStats* out) {
  NufftState& N = C.nu;
  hipStream_t st = C.stream;
  PLaunch L;
  int rc;
  if ((rc = persistent_config(C, N.MSP, {&N}, false, !P.single_trials, &L))) return rc;
  PArgs A;
  if ((rc = persistent_prepare(N, C.pw, P, L.nblk, L.lds_tables, C.persist_hier_barrier, L.pair, st, &A)))
    return rc;
  const char* trace_path = getenv("MFISTA_PERSISTENT_TRACE");
  const int tcap = 16384;
  if (trace_path && *trace_path) {
    if (C.trace.ensure((size_t)tcap * L.nblk * 16) || C.trace_tag.ensure((size_t)tcap * 4)) return MFISTA_E_HIP;
    MF_CHECK(hipMemsetAsync(C.trace.p, 0, (size_t)tcap * L.nblk * 16, st));  // unused entries stay 0
    A.trace = (unsigned long long*)C.trace.p;
    A.ttag = (int*)C.trace_tag.p;
    A.tcap = tcap;
  }
  if (C.pw.pargs.ensure(sizeof(PArgs))) return MFISTA_E_HIP;
  MF_CHECK(hipMemcpyAsync(C.pw.pargs.p, &A, sizeof(PArgs), hipMemcpyHostToDevice, st));
  const PArgs* dA = (const PArgs*)C.pw.pargs.p;
  void* args[] = {&dA};
  float kms = 0;
  if ((rc = timed_launch(L, args, L.nblk, st, C.no_alloc ? nullptr : &kms))) return rc;
  out->kernel_ms = kms;
  C.persist_blocks = L.nblk;
  if ((rc = persistent_collect(C.pw, P, st, out))) return rc;
  MF_CHECK(hipStreamSynchronize(st));
  if (A.trace) return read_trace(C, trace_path, tcap, L.nblk, *out, P, false);
  return 0;
}

// Launch configuration of the Toeplitz loop (toeplitz_body.h) for problems of
// the given geometries: [FP64 twiddle tables][stage scratch: two FP64
// sequences of the longer axis, double-buffered]
static int toeplitz_config(Context& C, const std::vector<const NufftState*>& sets, bool batch, PLaunch* L) {
  int dev = 0, ncu = 0, lds_max = 0;
  MF_CHECK(hipGetDevice(&dev));
  MF_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  MF_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t static_lds = 16 * kPW * 8 + sizeof(BlockState) + (kTD + 1) * 8 + sizeof(BarCtx) + 64;
  size_t sh = 0, tables = 0;
  for (const NufftState* N : sets) {
    const GridGeom& g = N->geom;
    // every stage's transform takes two buffers of one padded row or column
    // (T1 stages its z rows in the second one before its first FFT stage)
    const int Lm = g.Lx > g.Ly ? g.Lx : g.Ly;
    sh = std::max(sh, (size_t)2 * Lm * sizeof(double2));
    tables = std::max(tables, (toep_table_bytes(g.Lx, g.Ly) + 255) & ~(size_t)255);
  }
  L->lds_tables = (tables + sh + static_lds <= (size_t)lds_max) ? 1 : 0;
  if (L->lds_tables) sh += tables;
  if (sh + static_lds > (size_t)lds_max) return MFISTA_E_SIZE;
  L->sh = sh;
  if (batch)
    L->fn = L->lds_tables ? (const void*)k_mfista_toeplitz_batch<true> : (const void*)k_mfista_toeplitz_batch<false>;
  else
    L->fn = L->lds_tables ? (const void*)k_mfista_toeplitz<true> : (const void*)k_mfista_toeplitz<false>;
  int per_cu = 0;
  MF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, L->fn, MF_BLOCK, sh));
  if (per_cu < 1) return MFISTA_E_HIP;
  const int want = C.persist_blocks_per_cu > 0 ? C.persist_blocks_per_cu : 2;
  L->nblk = ncu * (per_cu < want ? per_cu : want);
  L->pair = 1;
  return 0;
}

// Allocates one Toeplitz-loop problem's workspace and fills its PArgs for nblk
// workgroups; op: its visibility set's operator data.  P.xinit and P.xout are
// device pointers.
static int toeplitz_prepare(const NufftState& N, const ToepOut& op, PersistWork& W, const RunParams& P, int nblk,
                            int lds_tables, int hier, hipStream_t st, PArgs* out) {
  const GridGeom& g = N.geom;
  const int NX = P.NX, NY = P.NY, n = NX * NY, nky = NY + 1;
  if (P.tv_mode || NX != g.NX || NY != g.NY) return MFISTA_E_ARG;
  const size_t t64 = (size_t)nky * NX * sizeof(double2);
  if (W.px.ensure((size_t)n * 8) || W.pz.ensure((size_t)n * 8) || W.T64.ensure(2 * t64) ||
      W.part.ensure((size_t)nblk * kPW * 8) || W.state.ensure(sizeof(DevState)) ||
      W.cost.ensure((size_t)(P.maxiter > 0 ? P.maxiter : 1) * 8) || W.bar.ensure(kBarBytes) ||
      W.tvec.ensure((size_t)4 * n * 8) || W.u64.ensure(2 * t64) || W.xrot.ensure((size_t)4 * n * 8))
    return MFISTA_E_HIP;
  if (P.xinit != W.px.p)
    MF_CHECK(hipMemcpyAsync(W.px.p, P.xinit, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  MF_CHECK(hipMemsetAsync(W.bar.p, 0, kBarBytes, st));
  PArgs& A = *out;
  memset(&A, 0, sizeof(A));
  A.g = g;
  A.Px = N.Px;
  A.Py = N.Py;
  A.twxd = (const double2*)N.twxd.p;
  A.twyd = (const double2*)N.twyd.p;
  A.T64 = (double2*)W.T64.p;
  A.T64b = (double2*)((char*)W.T64.p + t64);
  A.pair = 1;
  A.x = (double*)W.px.p;
  A.z = (double*)W.pz.p;
  A.box = P.box_flag ? P.box : nullptr;
  A.maxiter = P.maxiter;
  A.nonneg = P.nonneg;
  A.lambda_l1 = P.lambda_l1;
  A.lambda_tsv = P.lambda_tsv;
  A.eps_abs = P.eps_abs;
  A.rel_tol = P.rel_tol;
  A.eta_first = P.eta_first;
  A.cinit = P.cinit;
  A.part = (double*)W.part.p;
  A.n_row = (NX + 1) / 2;
  A.n_ky = nky;
  A.lds_tables = lds_tables;
  A.nblk = nblk;
  A.cost = (double*)W.cost.p;
  A.S = (DevState*)W.state.p;
  A.bar = (unsigned long long*)W.bar.p;
  A.hier_barrier = hier;
  A.kh = (const double*)op.kh.p;
  A.bimg = (const double*)op.bimg.p;
  for (int q = 0; q < 4; q++) A.tv[q] = (double*)W.tvec.p + (size_t)q * n;
  A.u64a = (double2*)W.u64.p;
  A.u64b = (double2*)((char*)W.u64.p + t64);
  A.wv2 = op.wv2;
  for (int q = 0; q < 4; q++) A.xb[q] = (double*)W.xrot.p + (size_t)q * n;
  A.Pxh = op.Pxh;
  A.twxh = (const double2*)op.twxh64.p;
  return 0;
}

// Toeplitz loop of a single solve (toeplitz_body.h); toeplitz_setup has run
// for the problem of C.nu into C.tw.out.  Paired trial rounds always.
int solve_nufft_toeplitz(Context& C, const RunParams& P, RunStats* out) {
  NufftState& N = C.nu;
  PersistWork& W = C.pw;
  hipStream_t st = C.stream;
  PLaunch L;
  int rc;
  if ((rc = toeplitz_config(C, {&N}, false, &L))) return rc;
  PArgs A;
  if ((rc = toeplitz_prepare(N, C.tw.out, W, P, L.nblk, L.lds_tables, C.persist_hier_barrier, st, &A))) return rc;
  const char* trace_path = getenv("MFISTA_PERSISTENT_TRACE");
  const int tcap = 16384;
  if (trace_path && *trace_path) {
    if (C.trace.ensure((size_t)tcap * L.nblk * 16) || C.trace_tag.ensure((size_t)tcap * 4)) return MFISTA_E_HIP;
    MF_CHECK(hipMemsetAsync(C.trace.p, 0, (size_t)tcap * L.nblk * 16, st));
    A.trace = (unsigned long long*)C.trace.p;
    A.ttag = (int*)C.trace_tag.p;
    A.tcap = tcap;
  }
  if (W.pargs.ensure(sizeof(PArgs))) return MFISTA_E_HIP;
  MF_CHECK(hipMemcpyAsync(W.pargs.p, &A, sizeof(PArgs), hipMemcpyHostToDevice, st));
  const PArgs* dA = (const PArgs*)W.pargs.p;
  void* args[] = {&dA};
  float kms = 0;
  if ((rc = timed_launch(L, args, L.nblk, st, C.no_alloc ? nullptr : &kms))) return rc;
  out->kernel_ms = kms;
  C.persist_blocks = L.nblk;
  if ((rc = persistent_collect(W, P, st, out))) return rc;
  MF_CHECK(hipStreamSynchronize(st));
  if (A.trace) return read_trace(C, trace_path, tcap, L.nblk, *out, P, true);
  return 0;
}

// Deals the launch's workgroups to problems p0 .. p0+pn-1 (deal.h): blk[b] =
// (problem, index within the problem), nb[i] = workgroups of problem i; the
// work per iteration of a problem is its pixels plus its gridding taps.
static int deal_workgroups(const std::vector<NufftState*>& sets, const std::vector<RunParams>& P, int p0, int pn,
                           int nblk, std::vector<int2>& blk, std::vector<int>& nb) {
  std::vector<double> w(pn);
  for (int i = 0; i < pn; i++)
    w[i] = (double)P[p0 + i].NX * P[p0 + i].NY + 4.0 * sets[p0 + i]->M * sets[p0 + i]->MSP;
  std::vector<std::pair<int, int>> d;
  if (!deal_workgroups(w, nblk, d, nb)) return MFISTA_E_ARG;
  blk.resize(d.size());
  for (size_t b = 0; b < d.size(); b++) blk[b] = make_int2(d[b].first, d[b].second);
  return 0;
}

// Launches problems p0 .. p0+pn-1 (already dealt) and collects their outcomes.
static int launch_dealt(Context& C, const PLaunch& L, std::vector<PArgs>& A, const std::vector<int2>& blk,
                        const std::vector<RunParams>& P, int p0, std::vector<RunStats>& out, float* kernel_ms) {
  const int pn = (int)A.size();
  int nlaunch = 0;
  for (int b = 0; b < L.nblk; b++)
    if (blk[b].x >= 0) nlaunch = b + 1;
  // workgroups past the last dealt one are not launched; undealt ones in
  // between return at once
  hipStream_t st = C.stream;
  if (C.batch_args.ensure((size_t)pn * sizeof(PArgs)) || C.batch_blk.ensure((size_t)nlaunch * sizeof(int2)))
    return MFISTA_E_HIP;
  MF_CHECK(hipMemcpyAsync(C.batch_args.p, A.data(), (size_t)pn * sizeof(PArgs), hipMemcpyHostToDevice, st));
  MF_CHECK(hipMemcpyAsync(C.batch_blk.p, blk.data(), (size_t)nlaunch * sizeof(int2), hipMemcpyHostToDevice, st));
  MF_CHECK(hipStreamSynchronize(st));
  const PArgs* dargs = (const PArgs*)C.batch_args.p;
  const int2* dblk = (const int2*)C.batch_blk.p;
  void* args[] = {&dargs, &dblk};
  float kms = 0;
  int rc;
  if ((rc = timed_launch(L, args, nlaunch, st, &kms))) return rc;
  *kernel_ms += kms;
  for (int i = 0; i < pn; i++) {
    out[p0 + i].kernel_ms = kms;
    if ((rc = persistent_collect(C.bwork[p0 + i], P[p0 + i], st, &out[p0 + i]))) return rc;
  }
  MF_CHECK(hipStreamSynchronize(st));
  return 0;
}

// Batched solves: problem i uses sets[i] (its NUFFT set-up) and C.bwork[i].
// Problems beyond one workgroup each go to further launches.
int solve_nufft_batch(Context& C, const std::vector<NufftState*>& sets, const std::vector<RunParams>& P,
                      std::vector<RunStats>& out, float* kernel_ms) {
  const int np = (int)P.size();
  hipStream_t st = C.stream;
  if (np == 0) return 0;
  const int MSP = sets[0]->MSP;
  std::vector<const NufftState*> cs(sets.begin(), sets.end());
  for (const NufftState* N : cs)
    if (N->MSP != MSP) return MFISTA_E_ARG;
  PLaunch L;
  int rc;
  if ((rc = persistent_config(C, MSP, cs, true, !P[0].single_trials, &L))) return rc;
  C.persist_blocks = L.nblk;
  if ((int)C.bwork.size() < np) C.bwork.resize(np);
  out.assign(np, RunStats());
  *kernel_ms = 0;
  for (int p0 = 0; p0 < np; p0 += L.nblk) {
    const int pn = std::min(np - p0, L.nblk);
    std::vector<int2> blk;
    std::vector<int> nb;
    if ((rc = deal_workgroups(sets, P, p0, pn, L.nblk, blk, nb))) return rc;
    std::vector<PArgs> A(pn);
    for (int i = 0; i < pn; i++)
      if ((rc = persistent_prepare(*sets[p0 + i], C.bwork[p0 + i], P[p0 + i], nb[i], L.lds_tables,
                                   C.persist_hier_barrier, L.pair, st, &A[i])))
        return rc;
    if ((rc = launch_dealt(C, L, A, blk, P, p0, out, kernel_ms))) return rc;
  }
  return 0;
}

// Independent Toeplitz-loop solves (L1 / L1+TSV, square images) in cooperative
// launches: problem i uses sets[i], ops[i] and C.bwork[i]; problems beyond one
// workgroup each go to further launches.
int solve_nufft_toeplitz_batch(Context& C, const std::vector<NufftState*>& sets,
                               const std::vector<const ToepOut*>& ops, const std::vector<RunParams>& P,
                               std::vector<RunStats>& out, float* kernel_ms) {
  const int np = (int)P.size();
  hipStream_t st = C.stream;
  if (np == 0) return 0;
  if ((int)ops.size() != np || (int)sets.size() != np) return MFISTA_E_ARG;
  std::vector<const NufftState*> cs(sets.begin(), sets.end());
  PLaunch L;
  int rc;
  if ((rc = toeplitz_config(C, cs, true, &L))) return rc;
  C.persist_blocks = L.nblk;
  if ((int)C.bwork.size() < np) C.bwork.resize(np);
  out.assign(np, RunStats());
  *kernel_ms = 0;
  for (int p0 = 0; p0 < np; p0 += L.nblk) {
    const int pn = std::min(np - p0, L.nblk);
    std::vector<int2> blk;
    std::vector<int> nb;
    if ((rc = deal_workgroups(sets, P, p0, pn, L.nblk, blk, nb))) return rc;
    std::vector<PArgs> A(pn);
    for (int i = 0; i < pn; i++)
      if ((rc = toeplitz_prepare(*sets[p0 + i], *ops[p0 + i], C.bwork[p0 + i], P[p0 + i], nb[i], L.lds_tables,
                                 C.persist_hier_barrier, st, &A[i])))
        return rc;
    if ((rc = launch_dealt(C, L, A, blk, P, p0, out, kernel_ms))) return rc;
  }
  return 0;
}

// A coupled movie: frame i uses sets[i] and C.bwork[i]; all frames run in one
// launch with one barrier and one control step (matlab/MFISTA_movie_eht.m).
// w[t] = 1 / (t_{t+1} - t_t), t = 0 .. nframes-2.
int solve_nufft_movie(Context& C, const std::vector<NufftState*>& sets, const std::vector<RunParams>& P,
                      double lambda_t, const std::vector<double>& w, std::vector<RunStats>& out,
                      float* kernel_ms) {
  const int nf = (int)P.size();
  hipStream_t st = C.stream;
  if (nf == 0 || (int)w.size() != nf - 1) return MFISTA_E_ARG;
  const int MSP = sets[0]->MSP;
  std::vector<const NufftState*> cs(sets.begin(), sets.end());
  for (const NufftState* N : cs)
    if (N->MSP != MSP) return MFISTA_E_ARG;
  for (const RunParams& R : P)
    if (R.tv_mode || R.NX != P[0].NX || R.NY != P[0].NY) return MFISTA_E_ARG;
  PLaunch L;
  int rc;
  if ((rc = persistent_config(C, MSP, cs, true, !P[0].single_trials, &L))) return rc;
  if (nf > L.nblk) return MFISTA_E_SIZE;  // every frame needs a workgroup of the one launch
  if ((int)C.bwork.size() < nf) C.bwork.resize(nf);
  out.assign(nf, RunStats());
  *kernel_ms = 0;
  std::vector<int2> blk;
  std::vector<int> nb;
  if ((rc = deal_workgroups(sets, P, 0, nf, L.nblk, blk, nb))) return rc;
  std::vector<PArgs> A(nf);
  for (int i = 0; i < nf; i++)
    if ((rc = persistent_prepare(*sets[i], C.bwork[i], P[i], nb[i], L.lds_tables, C.persist_hier_barrier, L.pair,
                                 st, &A[i])))
      return rc;
  int ngb = 0;
  for (int i = 0; i < nf; i++) ngb += nb[i];
  if (C.movie_part.ensure((size_t)ngb * kPW * 8)) return MFISTA_E_HIP;
  int g0 = 0;
  for (int i = 0; i < nf; i++) {
    PArgs& a = A[i];
    a.ngb = ngb;
    a.gb0 = g0;
    g0 += nb[i];
    a.part = (double*)C.movie_part.p;
    a.bar = A[0].bar;
    a.lambda_t = lambda_t;
    a.wt_prev = i > 0 ? w[i - 1] : 0.0;
    a.wt_next = i < nf - 1 ? w[i] : 0.0;
    a.z_prev = i > 0 ? A[i - 1].z : nullptr;
    a.z_next = i < nf - 1 ? A[i + 1].z : nullptr;
    a.x_next = i < nf - 1 ? A[i + 1].x : nullptr;
    a.xn_next[0] = i < nf - 1 ? A[i + 1].xn[0] : nullptr;
    a.xn_next[1] = i < nf - 1 ? A[i + 1].xn[1] : nullptr;
  }
  return launch_dealt(C, L, A, blk, P, 0, out, kernel_ms);
}

}  // namespace mf
