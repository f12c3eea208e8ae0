// autotune.hip -- the persistent solver's launch-configuration autotuner
// (see autotune.h): candidate generation, timing, the process-wide table and
// its text file.
#include "autotune.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>

#include "persistent_body.h"

namespace mf {

bool TuneKey::operator<(const TuneKey& o) const {
  const int a[6] = {NX, NY, preset, reg, m_log2, nprob_log2};
  const int b[6] = {o.NX, o.NY, o.preset, o.reg, o.m_log2, o.nprob_log2};
  for (int i = 0; i < 6; i++)
    if (a[i] != b[i]) return a[i] < b[i];
  return false;
}

static int floor_log2(int v) {
  int l = 0;
  while (v > 1) {
    v >>= 1;
    l++;
  }
  return l;
}

TuneKey tune_key(int NX, int NY, int preset, int reg, int M, int nprob) {
  TuneKey k;
  k.NX = NX;
  k.NY = NY;
  k.preset = preset;
  k.reg = reg;
  k.m_log2 = floor_log2(M > 0 ? M : 1);
  k.nprob_log2 = floor_log2(nprob > 0 ? nprob : 1);
  return k;
}

// the table (guarded by the library mutex, held by every entry point)
static std::map<TuneKey, TuneChoice> g_table;
static long g_runs = 0;
static std::string g_cache_path;

long tune_runs() { return g_runs; }

bool tune_lookup(const TuneKey& k, TuneChoice* out) {
  auto it = g_table.find(k);
  if (it == g_table.end()) return false;
  *out = it->second;
  return true;
}

static int save_table(const char* path);

void tune_store(const TuneKey& k, const TuneChoice& c) {
  g_table[k] = c;
  if (!g_cache_path.empty()) (void)save_table(g_cache_path.c_str());
}

void tune_apply(Context& C, const TuneChoice& c) {
  C.grid_tile_x = c.tile_x;
  C.grid_tile_y = c.tile_y;
  C.persist_blocks_per_cu = c.blocks_per_cu;
  C.persist_pair = c.paired;
  C.tune_engine = c.engine;
  C.fgp_tw = c.fgp_tw;
  C.tune_loop = c.loop;
}

void tune_restore(Context& C) {
  C.grid_tile_x = 16;
  C.grid_tile_y = 16;
  C.persist_blocks_per_cu = C.def_blocks_per_cu;
  C.persist_pair = C.def_pair;
  C.tune_engine = -1;
  C.fgp_tw = 0;
  C.tune_loop = 0;
}

std::vector<TuneChoice> tune_candidates(int reg, bool batch, bool toep) {
  static const int tiles[3][2] = {{16, 16}, {32, 16}, {32, 32}};
  std::vector<TuneChoice> v;
  static const int fgp_shapes[4] = {0, 64, 32, 16};
  for (const auto& t : tiles)
    for (int bpc = 1; bpc <= 2; bpc++)
      for (int pair = 1; pair >= (reg == 2 ? 1 : 0); pair--)  // TV rounds are never paired
        for (int f = 0; f < (reg == 2 ? 4 : 1); f++) {      // FGP shapes for TV only
          TuneChoice c;
          c.tile_x = t[0];
          c.tile_y = t[1];
          c.blocks_per_cu = bpc;
          c.paired = (reg == 2) ? 0 : pair;
          c.fgp_tw = fgp_shapes[f];
          if (toep) c.loop = MFISTA_LOOP_GRIDDING;
          v.push_back(c);
        }
  if (toep)  // the Toeplitz loop: workgroups per CU are its only knob (persistent launches)
    for (int bpc = 1; bpc <= 2; bpc++) {
      TuneChoice c;
      c.blocks_per_cu = bpc;
      c.loop = MFISTA_LOOP_TOEPLITZ;
      v.push_back(c);
    }
  if (batch && reg != 2) {  // the wide engine: the gridding tile is its only knob (gridding steps)
    for (const auto& t : tiles) {
      TuneChoice c;
      c.tile_x = t[0];
      c.tile_y = t[1];
      c.engine = MFISTA_ENGINE_STREAM;
      if (toep) c.loop = MFISTA_LOOP_GRIDDING;
      v.push_back(c);
    }
    if (toep) {  // Toeplitz steps
      TuneChoice c;
      c.engine = MFISTA_ENGINE_STREAM;
      c.loop = MFISTA_LOOP_TOEPLITZ;
      v.push_back(c);
    }
  }
  return v;
}

int tune_search(Context& C, int reg, bool batch, const std::function<int(double*)>& run, TuneChoice* best,
                std::vector<TuneChoice>* all, bool toep) {
  std::vector<TuneChoice> cand = tune_candidates(reg, batch, toep);
  int ibest = -1;
  for (size_t i = 0; i < cand.size(); i++) {
    TuneChoice& c = cand[i];
    c.us_per_iter = INFINITY;
    tune_apply(C, c);
    for (int rep = 0; rep < 2; rep++) {
      double us = 0;
      const int rc = run(&us);
      g_runs++;
      if (rc) {
        tune_restore(C);
        return rc;
      }
      if (us < c.us_per_iter) c.us_per_iter = us;
    }
    // the occupancy limit can make a blocks-per-CU candidate the same launch
    // as a smaller one; record the workgroups-per-CU the launch really used
    int dev = 0, ncu = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (c.engine == MFISTA_ENGINE_PERSISTENT && ncu > 0 && C.persist_blocks > 0)
      c.blocks_per_cu = C.persist_blocks / ncu;
    if (ibest < 0 || c.us_per_iter < cand[ibest].us_per_iter) ibest = (int)i;
  }
  tune_restore(C);
  if (ibest < 0) return MFISTA_E_ARG;
  *best = cand[ibest];
  if (all) *all = cand;
  return 0;
}

void to_public(const TuneKey& k, const TuneChoice& c, mfista_tune* t) {
  t->NX = k.NX;
  t->NY = k.NY;
  t->preset = k.preset;
  t->reg = k.reg;
  t->m_log2 = k.m_log2;
  t->nprob_log2 = k.nprob_log2;
  t->tile_x = c.tile_x;
  t->tile_y = c.tile_y;
  t->blocks_per_cu = c.blocks_per_cu;
  t->paired = c.paired;
  t->engine = c.engine;
  t->fgp_tw = c.fgp_tw;
  t->us_per_iter = c.us_per_iter;
  t->loop = c.loop;
}

// text file: a header line, then one entry per line,
// "NX NY preset reg m_log2 nprob_log2 tile_x tile_y blocks_per_cu paired engine fgp_tw loop us_per_iter"
// (lines without the loop field, 13 fields, load with loop 0)
static int save_table(const char* path) {
  FILE* f = fopen(path, "w");
  if (!f) return MFISTA_E_ARG;
  fprintf(f, "# mfista_amd tune table: NX NY preset reg m_log2 nprob_log2 tile_x tile_y blocks_per_cu paired "
             "engine fgp_tw loop us_per_iter\n");
  for (const auto& e : g_table) {
    const TuneKey& k = e.first;
    const TuneChoice& c = e.second;
    fprintf(f, "%d %d %d %d %d %d %d %d %d %d %d %d %d %.3f\n", k.NX, k.NY, k.preset, k.reg, k.m_log2,
            k.nprob_log2, c.tile_x, c.tile_y, c.blocks_per_cu, c.paired, c.engine, c.fgp_tw, c.loop,
            c.us_per_iter);
  }
  return fclose(f) == 0 ? 0 : MFISTA_E_ARG;
}

static int load_table(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return MFISTA_E_ARG;
  char line[512];
  int n = 0;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '#') continue;
    TuneKey k;
    TuneChoice c;
    double f13 = 0, f14 = 0;
    const int nf = sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d %d %lf %lf", &k.NX, &k.NY, &k.preset, &k.reg,
                          &k.m_log2, &k.nprob_log2, &c.tile_x, &c.tile_y, &c.blocks_per_cu, &c.paired, &c.engine,
                          &c.fgp_tw, &f13, &f14);
    if (nf == 14) {  // ... loop us_per_iter
      c.loop = (int)f13;
      if ((double)c.loop != f13) continue;
      c.us_per_iter = f14;
    } else if (nf == 13) {  // ... us_per_iter
      c.loop = 0;
      c.us_per_iter = f13;
    } else {
      continue;
    }
    if (!tile_supported(c.tile_x, c.tile_y) || c.blocks_per_cu < 1 || c.blocks_per_cu > 4 ||
        (c.engine != MFISTA_ENGINE_PERSISTENT && c.engine != MFISTA_ENGINE_STREAM) ||
        (c.fgp_tw != 0 && c.fgp_tw != 16 && c.fgp_tw != 32 && c.fgp_tw != 64) ||
        (c.loop != 0 && c.loop != MFISTA_LOOP_GRIDDING && c.loop != MFISTA_LOOP_TOEPLITZ))
      continue;
    g_table[k] = c;
    n++;
  }
  fclose(f);
  return n;
}

void tune_load_env() {
  static bool done = false;
  if (done) return;
  done = true;
  if (const char* p = getenv("MFISTA_TUNE_CACHE")) {
    g_cache_path = p;
    (void)load_table(p);
  }
}

}  // namespace mf

using namespace mf;

extern "C" {

int mfista_tune_table(mfista_tune* out, int cap) {
  std::lock_guard<std::mutex> lock(global_mutex());
  int n = 0;
  for (const auto& e : g_table) {
    if (out && n < cap) to_public(e.first, e.second, &out[n]);
    n++;
  }
  return n;
}

int mfista_tune_save(const char* path) {
  std::lock_guard<std::mutex> lock(global_mutex());
  if (!path) return MFISTA_E_ARG;
  return save_table(path);
}

int mfista_tune_load(const char* path) {
  std::lock_guard<std::mutex> lock(global_mutex());
  if (!path) return MFISTA_E_ARG;
  return load_table(path);
}

void mfista_tune_clear(void) {
  std::lock_guard<std::mutex> lock(global_mutex());
  g_table.clear();
}

long mfista_tune_runs(void) {
  std::lock_guard<std::mutex> lock(global_mutex());
  return g_runs;
}

}  // extern "C"
