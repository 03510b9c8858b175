// batch.hpp -- many independent scenarios advanced together on one GPU.
//
// Each scenario is a ShallowWaterEngine (its own device state, tables and step state
// machine).  A batch step plans every scenario's step on the host, writes the parameter
// blocks of all scenarios into one device table, and issues one launch per kernel over
// a work-list concatenated across scenarios: explicit tiles, line-solve blocks (one
// launch per line-length class), boundary cells.  Blocks of K steps are captured as one
// hipGraph per batch composition and replayed with that block's parameter table, so
// scenarios at different points of their schedules (bootstrap, steady state) share a
// graph.  Results equal those of each scenario run alone.
#pragma once

#include <map>
#include <memory>
#include <set>
#include <vector>

#include "engine.hpp"

namespace cel {

hipError_t launch_explicit_batched(int rows, const ExplicitJob *jobs, const int *starts, int njobs, int total_tiles,
                                   hipStream_t st);
hipError_t launch_solve_batched(int shape, const SolveJob *jobs, const int *starts, int njobs, int total_blocks,
                                hipStream_t st);
hipError_t launch_boundary_batched(const BoundaryJob *jobs, const int *starts, int njobs, int total_cells,
                                   int mirror_pass, hipStream_t st);

class Batch {
public:
    Batch();
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    int add(const cel_settings &s, const cel_boundary bnd[4], const float *bottom3, const float *state3);
    // takes over an engine created elsewhere (e.g. on another host thread while the batch
    // runs); e is moved from only on success
    int adopt(std::unique_ptr<ShallowWaterEngine> &e);
    void remove(int id);
    void setPaused(int id, bool paused);     // paused scenarios are skipped by step()
    void step(int nsteps);
    void setGraphSteps(int k);
    int autotune(int repeats, bool tune_lines);
    void sync();
    float timedSteps(int nsteps);
    ShallowWaterEngine &scenario(int id);      // synchronises the batch stream first
    void setState(int id, const float *state3);   // hot start of one member (clears its sentinel)
    // members whose steps produced a non-finite value since they were added or last given a
    // state (waits for the stream)
    std::vector<int> failed();
    int count() const { return (int)scen_.size(); }
    size_t deviceBytes() const;
    void ids(int *out) const;

private:
    struct Composition {
        std::vector<int> ids;                  // scenario order of the explicit/boundary lists
        std::vector<int> solve_order;          // 2 i (x-lines) / 2 i + 1 (y-lines), sorted by block shape
        std::vector<int> sig;                  // per-scenario work counts and classes
        int total_tiles = 0, total_cells = 0;
        bool one_pass = true;                  // every scenario's boundaries fit one launch (pass 2)
        int cls_first[kNumShapes] = {}, cls_count[kNumShapes] = {}, cls_blocks[kNumShapes] = {};
        int cls_starts_off[kNumShapes] = {};   // offsets into the starts buffer (ints)
        int exp_starts_off = 0, bnd_starts_off = 0;
    };
    void run(int m, bool use_graph);
    // plans m steps of every scenario (m == 1 when not advancing), brings the composition
    // and the parameter table up to date and fills the host table
    void prepare(int m, bool advance, std::vector<StepLaunch> &plans, std::vector<ShallowWaterEngine *> &eng);
    void fillTable(const std::vector<StepLaunch> &plans, const std::vector<ShallowWaterEngine *> &eng, int m, char *h);
    void buildComposition(const std::vector<StepLaunch> &first, const std::vector<int> &ids);
    void tuneSharedLineShape(int repeats);
    void issue(const char *blob, size_t stride, int m);
    void clearGraphs();
    void check(hipError_t e, const char *what) const;
    void ensureBlob(size_t bytes);

    hipStream_t stream_ = nullptr;
    std::map<int, std::unique_ptr<ShallowWaterEngine>> scen_;
    // numeric sentinel words, one slot per member (ExplicitParams::status of its jobs)
    static constexpr int kMaxMembers = 4096;
    int *d_status_ = nullptr;
    std::map<int, int> slot_;                  // member id -> sentinel slot
    std::vector<int> free_slots_;
    std::set<int> paused_;
    int next_id_ = 1;
    int graph_steps_ = 0;
    int rows_ = kDefaultBandRows;              // band rows of the batched explicit launch
    Composition comp_;
    bool comp_valid_ = false;
    int *d_starts_ = nullptr;
    size_t starts_cap_ = 0;
    // per-step parameter table: device copy and two pinned host buffers
    char *d_blob_ = nullptr;
    char *h_blob_[2] = {};
    size_t blob_cap_ = 0;
    hipEvent_t blob_ev_[2] = {};
    bool blob_used_[2] = {};
    int blob_slot_ = 0;
    size_t sec_exp_ = 0, sec_solve_ = 0, sec_bnd_ = 0, stride_ = 0;   // blob layout of one step
    struct GraphEntry { int m; hipGraphExec_t exec; };
    std::vector<GraphEntry> graphs_;
    hipEvent_t ev_[2] = {};
};

}  // namespace cel
