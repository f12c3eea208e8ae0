// autotune.h -- launch-configuration autotuner of the persistent NUFFT solver.
//
// The reference has one fixed schedule (c++/mfista_nufft_lib.cpp:643-678,
// configure_tile, balances gridding taps per OpenMP thread by rule).  Here the
// knobs whose best value depends on the problem's shape are timed on the GPU:
//   * the gridding tile: 16 x 16, 32 x 16 or 32 x 32 cells (1, 2 or 4 cells per
//     thread of the gatherer; larger tiles evaluate more empty cells per
//     visibility but have fewer shared edges flushed with atomics);
//   * workgroups per CU of the cooperative launch (1 .. the occupancy limit):
//     more workgroups fill stages with many tasks, fewer make every grid
//     barrier cheaper;
//   * paired backtracking rounds (the trials c and c*eta evaluated together)
//     or one trial per round;
//   * for batches, the engine: one persistent launch for the whole batch, or
//     the wide engine (every stage one launch for all problems, wide.hip);
//   * for L1+TV, the FGP block shape: pixels dealt one per thread along the
//     flat image, or in tiles of 64x4, 32x8 or 16x16 pixels per workgroup;
//   * for single solves where the Toeplitz loop applies, the loop: the
//     gridding loop's candidates above, or the Toeplitz loop at 1 or 2
//     workgroups per CU (its only knob).
// Every candidate solves the caller's own problem for a fixed number of
// iterations from its own start (no stopping rule); the fastest per
// iteration is kept in a process-wide table under the problem's shape key,
// and later solves of that shape use it.  The table can be saved to and
// loaded from a text file (MFISTA_TUNE_CACHE names one that is loaded when
// the first context is created and rewritten after every new entry).
#pragma once
#include <functional>
#include <vector>

#include "mfista.h"
#include "solver.h"

namespace mf {

struct TuneKey {
  int NX = 0, NY = 0, preset = 0, reg = 0, m_log2 = 0, nprob_log2 = 0;
  bool operator<(const TuneKey& o) const;
};

struct TuneChoice {
  int tile_x = 16, tile_y = 16, blocks_per_cu = 2, paired = 1;
  int engine = MFISTA_ENGINE_PERSISTENT;  // batches: MFISTA_ENGINE_STREAM = the wide engine
  int fgp_tw = 0;                         // L1+TV: FGP block shape (0 = along the flat image)
  int loop = 0;  // single solves: MFISTA_LOOP_GRIDDING or MFISTA_LOOP_TOEPLITZ (0: the loop by rule)
  double us_per_iter = 0;
};

// regulariser class of a problem: 0 L1, 1 L1+TSV, 2 L1+TV
inline int tune_reg(double lambda_tv, double lambda_tsv) { return lambda_tv > 0 ? 2 : (lambda_tsv > 0 ? 1 : 0); }
TuneKey tune_key(int NX, int NY, int preset, int reg, int M, int nprob);

bool tune_lookup(const TuneKey& k, TuneChoice* out);
void tune_store(const TuneKey& k, const TuneChoice& c);
// sets the context's knobs for the next set-up and launch; tune_restore puts
// back the defaults (environment overrides included)
void tune_apply(Context& C, const TuneChoice& c);
void tune_restore(Context& C);
// the candidates for a solve (batch = false) or a batch of this regulariser
// class; toep: a single solve the Toeplitz loop applies to (its candidates
// are added and the gridding loop's are marked as such)
std::vector<TuneChoice> tune_candidates(int reg, bool batch, bool toep = false);
// Times `run` (which solves with the context's current knobs and returns 0
// and the time per iteration of the launch, in microseconds) for every
// candidate, twice each; the fastest in *best, every candidate's time in
// *all (optional).  The context's knobs are restored afterwards.
int tune_search(Context& C, int reg, bool batch, const std::function<int(double*)>& run, TuneChoice* best,
                std::vector<TuneChoice>* all, bool toep = false);
long tune_runs();  // candidate solves timed so far in this process
void tune_load_env();

void to_public(const TuneKey& k, const TuneChoice& c, mfista_tune* t);

}  // namespace mf
