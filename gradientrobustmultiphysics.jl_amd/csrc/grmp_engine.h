// Batched Krylov engine of libgrmp_hip.so.
//
// Many independent sparse systems ("problems") are solved together by one
// sequence of kernel launches: every kernel walks a table of row tiles, each
// tile belonging to one problem, and a per-problem phase machine decides what
// a tile does (FP64 residual, FP32 inner CG step, FP64 update).  Each tile
// writes its partial dot products to its own slot; a separate reduce kernel
// (k_eng_reduce) then sums a problem's slots in tile order, so a problem's
// iterates do not depend on which other problems share the launch.  The launch sequence of one step never changes,
// so it is captured once as a hipGraph and replayed; problems join and leave
// between replays by rewriting the tile table.
#pragma once
#include "grmp_internal.h"

enum KPhase {
    PH_IDLE = 0,
    PH_RESID = 1,
    PH_INIT = 2,
    PH_ITER = 3,
    PH_UPDATE = 4,
    PH_DONE = 5,
    PH_FAIL = 6,
    PH_REINIT = 7,
    PH_INIT64 = 8  // start (or restart) FP64 CG from the FP64 residual
};

// Rows per engine tile (the unit of work of one block) and the largest grid.
#define ENGINE_TILE_ROWS 2048
#define ENGINE_MAX_GRID 4096

struct KProb {
    // matrix (CSR, 0-based), fixed-dof mask, Jacobi inverse diagonal
    const int32_t *rowptr;
    const int32_t *cols;
    const double *val64;
    const float *val32;
    const float *dinv32;
    const double *dinv64;
    const uint8_t *fixed;
    // FP64 outer vectors
    const double *b;
    double *x;
    double *r64;
    // FP32 inner vectors
    float *r, *p, *q, *dx;
    int32_t n, pad0_;
    int32_t phase, status;
    int32_t outer;      // FP64 residual refreshes so far
    int32_t max_outer;
    int32_t restart;    // restart the search direction at the next refresh
    int32_t discard;    // the FP32 sweep overflowed: drop its correction
    int32_t stall;      // consecutive refreshes that did not halve the best residual
    int32_t fp64;       // FP64 fallback: the FP32 sweeps stalled, plain FP64 CG finishes
    int32_t slow;       // the current FP32 sweep ran past its iteration cap
    int32_t fp64_after; // > 0: switch to FP64 when the first FP32 sweep took more iterations
    int32_t pad2_;
    int64_t sweep0;     // iterations done when the current FP32 sweep started
    double *p64, *q64;  // FP64 direction and product (alias the FP32 vector storage)
    int64_t iters, maxit;
    double bb, bnorm;   // ||b - A x0||^2 and its root (free rows)
    double rtol;        // target relative FP64 residual
    double delta;       // reliable-update threshold on the FP32 residual
    double scale;       // FP32 quantities are FP64 ones times scale
    double pscale;      // search-direction rescale at a refresh
    double rz, alpha, beta, relres, rmax;
    double best;        // smallest FP64 relative residual so far
};

struct KTile {
    int32_t prob, row0, row1, pad;
};

struct KRange {
    int32_t t0, nt;
};

struct EngineSlot {
    bool used = false;
    int32_t ntiles = 0;
};

class KEngine {
  public:
    KEngine(grmp_ctx *ctx, hipStream_t st, int cap_probs, int cap_tiles, int tile_rows);
    ~KEngine();
    int init();
    // registers a problem (device pointers valid until remove); returns slot
    int add(const KProb &p, int *slot);
    void remove(int slot);
    int nactive() const { return nact_; }
    // uploads the tile table if it changed (stream ordered)
    int sync_tables();
    // runs `steps` engine steps (graph replay when use_graph)
    int run(int steps, bool use_graph);
    // reads phase/status/iterations of all slots (blocking)
    int poll(std::vector<KProb> &out);
    int64_t total_tiles() const { return ttot_; }
    // SpMV kernel accounting: events around every engine SpMV launch when
    // profiling, and a device counter of algorithmic bytes (always on)
    void set_profile(bool on) { profile_ = on; }
    int stats(double *seconds, int64_t *launches, double *bytes, int64_t *steps);

  private:
    int capture(int grid);
    int launch_step(int grid);
    void launch_kernels(int grid, bool timed);
    grmp_ctx *ctx_;
    hipStream_t st_;
    int cap_p_, cap_t_, tile_rows_;
    KProb *dP_ = nullptr;
    KTile *dT_ = nullptr;
    KRange *dR_ = nullptr;
    double *dpart_ = nullptr;
    int *dnt_ = nullptr;
    KProb *hP_ = nullptr;       // pinned poll buffer
    KTile *hT_ = nullptr;       // pinned table staging
    KRange *hR_ = nullptr;
    std::vector<EngineSlot> slots_;
    std::vector<KProb> init_;
    int nact_ = 0;
    int64_t ttot_ = 0;
    bool dirty_ = true;
    int graph_grid_ = 0;
    hipGraphExec_t exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    bool profile_ = false;
    static constexpr int EV_PAIRS = 32;
    hipEvent_t ev_[2 * EV_PAIRS] = {};
    int npending_ = 0;
    unsigned long long *dbytes_ = nullptr;
    unsigned long long *dbytes_blk_ = nullptr;  // per-block SpMV bytes of the last step
    double spmv_s_ = 0.0;
    int64_t spmv_n_ = 0, steps_ = 0;
    int read_events(int pairs);
};

