// batch.cpp -- batched scenarios: one launch per kernel per step over all scenarios.
#include "batch.hpp"

#include <algorithm>
#include <cstring>

namespace cel {

namespace {
size_t align256(size_t v) { return (v + 255) / 256 * 256; }
}  // namespace

void Batch::check(hipError_t e, const char *what) const
{
    if (e != hipSuccess) throw Error(e == hipErrorOutOfMemory ? CEL_ENOMEM : CEL_EDEVICE,
                                     std::string(what) + ": " + hipGetErrorString(e));
}

Batch::Batch()
{
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    for (auto &e : blob_ev_) check(hipEventCreate(&e), "hipEventCreate");
    for (auto &e : ev_) check(hipEventCreate(&e), "hipEventCreate");
    void *mem = nullptr;
    check(hipMalloc(&mem, kMaxMembers * sizeof(int)), "hipMalloc(batch status)");
    d_status_ = static_cast<int *>(mem);
    check(hipMemsetAsync(d_status_, 0, kMaxMembers * sizeof(int), stream_), "clear batch status");
    for (int i = kMaxMembers - 1; i >= 0; --i) free_slots_.push_back(i);
}

Batch::~Batch()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    clearGraphs();
    scen_.clear();
    if (d_blob_) (void)hipFree(d_blob_);
    for (auto &h : h_blob_) if (h) (void)hipHostFree(h);
    if (d_starts_) (void)hipFree(d_starts_);
    if (d_status_) (void)hipFree(d_status_);
    for (auto &e : blob_ev_) if (e) (void)hipEventDestroy(e);
    for (auto &e : ev_) if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int Batch::add(const cel_settings &s, const cel_boundary bnd[4], const float *bottom3, const float *state3)
{
    if (s.correction_steps_num != 0) throw Error(CEL_EINVAL, "batched scenarios do not take corrector steps");
    if (s.state_precision == CEL_STATE_FP64) throw Error(CEL_EINVAL, "batched scenarios run in FP32 (plain or compensated)");
    if (free_slots_.empty()) throw Error(CEL_ENOMEM, "a batch holds at most 4096 scenarios");
    // the new engine is set up on its own stream while the batch's launches run: nothing
    // in flight touches it, and the next step rebuilds the work-lists
    auto e = std::make_unique<ShallowWaterEngine>(s, bnd, bottom3, state3);
    return adopt(e);
}

int Batch::adopt(std::unique_ptr<ShallowWaterEngine> &e)
{
    if (!e) throw Error(CEL_EINVAL, "null engine");
    if (e->settings().correction_steps_num != 0) throw Error(CEL_EINVAL, "batched scenarios do not take corrector steps");
    if (free_slots_.empty()) throw Error(CEL_ENOMEM, "a batch holds at most 4096 scenarios");
    e->sync();                                   // its uploads are complete before any batch step
    const int id = next_id_++;
    const int slot = free_slots_.back();
    check(hipMemsetAsync(d_status_ + slot, 0, sizeof(int), stream_), "clear member status");   // ordered before its steps
    free_slots_.pop_back();
    slot_[id] = slot;
    scen_[id] = std::move(e);
    comp_valid_ = false;
    return id;
}

void Batch::setState(int id, const float *state3)
{
    ShallowWaterEngine &e = scenario(id);
    e.setState(state3);
    check(hipMemsetAsync(d_status_ + slot_.at(id), 0, sizeof(int), stream_), "clear member status");
    sync();
}

std::vector<int> Batch::failed()
{
    std::vector<int> flags(kMaxMembers);
    check(hipMemcpyAsync(flags.data(), d_status_, kMaxMembers * sizeof(int), hipMemcpyDeviceToHost, stream_),
          "read batch status");
    sync();
    std::vector<int> out;
    for (const auto &kv : slot_)
        if (flags[kv.second] != 0) out.push_back(kv.first);
    return out;
}

void Batch::remove(int id)
{
    auto it = scen_.find(id);
    if (it == scen_.end()) throw Error(CEL_EINVAL, "no such scenario");
    sync();
    scen_.erase(it);
    free_slots_.push_back(slot_.at(id));
    slot_.erase(id);
    paused_.erase(id);
    comp_valid_ = false;
}

void Batch::setPaused(int id, bool paused)
{
    if (!scen_.count(id)) throw Error(CEL_EINVAL, "no such scenario");
    if (paused == (paused_.count(id) != 0)) return;
    sync();
    if (paused) paused_.insert(id);
    else paused_.erase(id);
    comp_valid_ = false;
}

ShallowWaterEngine &Batch::scenario(int id)
{
    auto it = scen_.find(id);
    if (it == scen_.end()) throw Error(CEL_EINVAL, "no such scenario");
    sync();
    return *it->second;
}

void Batch::ids(int *out) const
{
    int i = 0;
    for (const auto &kv : scen_) out[i++] = kv.first;
}

size_t Batch::deviceBytes() const
{
    size_t b = blob_cap_ + starts_cap_ * sizeof(int);
    for (const auto &kv : scen_) b += kv.second->deviceBytes();
    return b;
}

void Batch::sync() { check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

void Batch::clearGraphs()
{
    for (auto &g : graphs_) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
}

void Batch::setGraphSteps(int k)
{
    if (k < 0 || k > kMaxGraphSteps) throw Error(CEL_EINVAL, "steps per graph must lie in [0, 1024]");
    sync();
    clearGraphs();
    graph_steps_ = k;
}

void Batch::ensureBlob(size_t bytes)
{
    if (bytes <= blob_cap_) return;
    sync();
    clearGraphs();   // graphs hold the old table's address
    if (d_blob_) (void)hipFree(d_blob_);
    for (auto &h : h_blob_) if (h) (void)hipHostFree(h);
    d_blob_ = nullptr;
    h_blob_[0] = h_blob_[1] = nullptr;
    const size_t cap = std::max(bytes, blob_cap_ * 2);
    void *mem = nullptr;
    check(hipMalloc(&mem, cap), "hipMalloc(batch table)");
    d_blob_ = static_cast<char *>(mem);
    for (auto &h : h_blob_) {
        check(hipHostMalloc(&mem, cap, hipHostMallocDefault), "hipHostMalloc(batch table)");
        h = static_cast<char *>(mem);
    }
    blob_cap_ = cap;
    blob_used_[0] = blob_used_[1] = false;
}

// Work counts of each scenario (from its first planned step; they depend on the grid,
// the boundary configuration and the chosen shapes only), start offsets of every
// work-list.  The line solves are listed per line direction (x-lines carry the
// irregular-wave lines), grouped by block shape.
void Batch::buildComposition(const std::vector<StepLaunch> &first, const std::vector<int> &ids)
{
    const int ns = (int)ids.size();
    Composition c;
    c.ids = ids;
    std::vector<int> tiles(ns), cells(ns), shp(2 * ns), blocks(2 * ns);
    for (int i = 0; i < ns; ++i) {
        const StepLaunch &L = first[i];
        tiles[i] = L.ep.nstrips * L.ep.nbands;
        shp[2 * i] = L.shape_x;
        shp[2 * i + 1] = L.shape_y;
        int irr_lines = 0;
        for (int s = 0; s < 4; ++s) irr_lines += L.irr.count[s];
        const int W = kShapeW[L.shape_x];
        blocks[2 * i] = L.lp.nlines_x + (irr_lines + W - 1) / W;
        blocks[2 * i + 1] = L.lp.nlines_y;
        cells[i] = L.bp.ncells;
        c.sig.insert(c.sig.end(), {ids[i], tiles[i], shp[2 * i], shp[2 * i + 1], blocks[2 * i], blocks[2 * i + 1],
                                   cells[i]});
    }
    c.solve_order.resize(2 * ns);
    for (int i = 0; i < 2 * ns; ++i) c.solve_order[i] = i;
    std::stable_sort(c.solve_order.begin(), c.solve_order.end(), [&](int a, int b) { return shp[a] < shp[b]; });
    // starts buffer: explicit (ns+1), per shape (n_c+1), boundary (ns+1)
    std::vector<int> st;
    c.exp_starts_off = 0;
    st.push_back(0);
    for (int i = 0; i < ns; ++i) st.push_back(st.back() + tiles[i]);
    c.total_tiles = st.back();
    int pos = 0;
    for (int k = 0; k < kNumShapes; ++k) {
        c.cls_first[k] = pos;
        c.cls_starts_off[k] = (int)st.size();
        st.push_back(0);
        while (pos < 2 * ns && shp[c.solve_order[pos]] == k) {
            st.push_back(st.back() + blocks[c.solve_order[pos]]);
            ++pos;
        }
        c.cls_count[k] = pos - c.cls_first[k];
        c.cls_blocks[k] = st.back();
    }
    c.bnd_starts_off = (int)st.size();
    st.push_back(0);
    for (int i = 0; i < ns; ++i) st.push_back(st.back() + cells[i]);
    c.total_cells = st.back();
    c.one_pass = true;
    for (const StepLaunch &L : first) c.one_pass = c.one_pass && L.bp.mirror_pass == 2;
    sync();
    clearGraphs();
    if (st.size() > starts_cap_) {
        if (d_starts_) (void)hipFree(d_starts_);
        void *mem = nullptr;
        check(hipMalloc(&mem, st.size() * sizeof(int) * 2), "hipMalloc(batch lists)");
        d_starts_ = static_cast<int *>(mem);
        starts_cap_ = st.size() * 2;
    }
    check(hipMemcpy(d_starts_, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice), "upload batch lists");
    sec_exp_ = 0;
    sec_solve_ = align256((size_t)ns * sizeof(ExplicitJob));
    sec_bnd_ = sec_solve_ + align256((size_t)2 * ns * sizeof(SolveJob));
    stride_ = sec_bnd_ + align256((size_t)ns * sizeof(BoundaryJob));
    comp_ = std::move(c);
    comp_valid_ = true;
}

// Launches of m steps whose parameter tables sit at blob + k * stride.
void Batch::issue(const char *blob, size_t stride, int m)
{
    const int ns = (int)comp_.ids.size();
    for (int k = 0; k < m; ++k) {
        const char *b = blob + (size_t)k * stride;
        const ExplicitJob *ej = reinterpret_cast<const ExplicitJob *>(b + sec_exp_);
        const SolveJob *sj = reinterpret_cast<const SolveJob *>(b + sec_solve_);
        const BoundaryJob *bj = reinterpret_cast<const BoundaryJob *>(b + sec_bnd_);
        check(launch_explicit_batched(rows_, ej, d_starts_ + comp_.exp_starts_off, ns, comp_.total_tiles, stream_),
              "explicit_batched");
        for (int c = 0; c < kNumShapes; ++c)
            if (comp_.cls_count[c] > 0)
                check(launch_solve_batched(c, sj + comp_.cls_first[c], d_starts_ + comp_.cls_starts_off[c],
                                           comp_.cls_count[c], comp_.cls_blocks[c], stream_), "solve_batched");
        if (comp_.one_pass)
            check(launch_boundary_batched(bj, d_starts_ + comp_.bnd_starts_off, ns, comp_.total_cells, 2, stream_),
                  "boundary_batched");
        else
            for (int pass = 1; pass >= 0; --pass)
                check(launch_boundary_batched(bj, d_starts_ + comp_.bnd_starts_off, ns, comp_.total_cells, pass,
                                              stream_), "boundary_batched");
    }
}

void Batch::prepare(int m, bool advance, std::vector<StepLaunch> &plans, std::vector<ShallowWaterEngine *> &eng)
{
    std::vector<int> ids;
    eng.clear();
    for (auto &kv : scen_)
        if (!paused_.count(kv.first)) { ids.push_back(kv.first); eng.push_back(kv.second.get()); }
    const int ns = (int)ids.size();
    plans.assign((size_t)m * ns, StepLaunch{});
    for (int k = 0; k < m; ++k)
        for (int i = 0; i < ns; ++i) {
            StepLaunch &L = plans[(size_t)k * ns + i];
            eng[i]->planStep(L, advance);
            L.ep.nbands = (L.ep.ny + rows_ - 1) / rows_;      // the batch's band height
        }
    for (const auto &L : plans)
        if (L.scheme == CEL_CORRECTOR || L.ncopy > 0) throw Error(CEL_ESTATE, "corrector step in a batch");
    std::vector<int> sig;
    if (comp_valid_)
        for (int i = 0; i < ns; ++i) {
            const StepLaunch &L = plans[i];
            int irr_lines = 0;
            for (int s = 0; s < 4; ++s) irr_lines += L.irr.count[s];
            const int W = kShapeW[L.shape_x];
            sig.insert(sig.end(), {ids[i], L.ep.nstrips * L.ep.nbands, L.shape_x, L.shape_y,
                                   L.lp.nlines_x + (irr_lines + W - 1) / W, L.lp.nlines_y, L.bp.ncells});
        }
    if (!comp_valid_ || sig != comp_.sig) {
        std::vector<StepLaunch> first(plans.begin(), plans.begin() + ns);
        buildComposition(first, ids);
    }
    ensureBlob(stride_ * (size_t)std::max(m, std::max(graph_steps_, 1)));
}

// Host table: per step, explicit jobs in scenario order, solve jobs in shape order,
// boundary jobs in scenario order.
void Batch::fillTable(const std::vector<StepLaunch> &plans, const std::vector<ShallowWaterEngine *> &eng, int m,
                      char *h)
{
    const int ns = (int)eng.size();
    for (int k = 0; k < m; ++k) {
        char *b = h + (size_t)k * stride_;
        ExplicitJob *ej = reinterpret_cast<ExplicitJob *>(b + sec_exp_);
        SolveJob *sj = reinterpret_cast<SolveJob *>(b + sec_solve_);
        BoundaryJob *bj = reinterpret_cast<BoundaryJob *>(b + sec_bnd_);
        for (int i = 0; i < ns; ++i) {
            const StepLaunch &L = plans[(size_t)k * ns + i];
            ej[i].p = L.ep;
            ej[i].p.status = d_status_ + slot_.at(comp_.ids[i]);
            ej[i].scheme = L.scheme;
            ej[i].comp = eng[i]->compensated();
            ej[i].fast = eng[i]->fastArith();
            bj[i].p = L.bp;
            bj[i].p.clock = nullptr;
        }
        for (int q = 0; q < 2 * ns; ++q) {
            const int e = comp_.solve_order[q], i = e / 2;
            const StepLaunch &L = plans[(size_t)k * ns + i];
            sj[q].p = L.lp;
            sj[q].irr = L.irr;
            sj[q].irr.clock = nullptr;
            if (e % 2 == 0) {
                sj[q].p.nlines_y = 0;                             // x-lines and irregular-wave lines
            } else {
                sj[q].p.nlines_x = 0;                             // y-lines
                for (int s = 0; s < 4; ++s) sj[q].irr.count[s] = 0;
            }
            sj[q].comp = eng[i]->compensated();
        }
    }
}

void Batch::run(int m, bool use_graph)
{
    if (scen_.size() == paused_.size() || m == 0) return;    // nothing to advance
    std::vector<StepLaunch> plans;
    std::vector<ShallowWaterEngine *> eng;
    prepare(m, true, plans, eng);
    char *h = h_blob_[blob_slot_];
    if (blob_used_[blob_slot_]) check(hipEventSynchronize(blob_ev_[blob_slot_]), "table event");
    fillTable(plans, eng, m, h);
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        for (const auto &g : graphs_)
            if (g.m == m) exec = g.exec;
        if (!exec) {
            hipGraph_t graph = nullptr;
            // relaxed: other host threads may be creating engines (allocations, copies on their
            // own streams) while this stream is captured
            check(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
            try {
                issue(d_blob_, stride_, m);
            } catch (...) {
                (void)hipStreamEndCapture(stream_, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            check(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            check(e, "hipGraphInstantiate");
            graphs_.push_back(GraphEntry{m, exec});
        }
    }
    check(hipMemcpyAsync(d_blob_, h, stride_ * (size_t)m, hipMemcpyHostToDevice, stream_), "upload batch table");
    check(hipEventRecord(blob_ev_[blob_slot_], stream_), "table event");
    blob_used_[blob_slot_] = true;
    blob_slot_ ^= 1;
    if (exec) check(hipGraphLaunch(exec, stream_), "hipGraphLaunch");
    else issue(d_blob_, stride_, m);
}

// Band height of the batched explicit launch: each candidate timed on the current step's
// parameters (the launch writes only buffers the next step rewrites).
int Batch::autotune(int repeats, bool tune_lines)
{
    if (repeats < 1 || repeats > 1000) throw Error(CEL_EINVAL, "repeats must lie in [1, 1000]");
    sync();
    if (tune_lines)
        for (auto &kv : scen_) kv.second->autotune(repeats, nullptr);
    if (scen_.size() == paused_.size()) return rows_;
    const int cand[3] = {4, 8, 16};
    float cand_ms[3] = {-1.f, -1.f, -1.f};
    for (int round = 0; round < 3; ++round)     // interleaved rounds, fastest round counts
        for (int c = 0; c < 3; ++c) {
            rows_ = cand[c];
            std::vector<StepLaunch> plans;
            std::vector<ShallowWaterEngine *> eng;
            prepare(1, false, plans, eng);
            sync();
            char *h = h_blob_[0];
            fillTable(plans, eng, 1, h);
            check(hipMemcpy(d_blob_, h, stride_, hipMemcpyHostToDevice), "upload batch table");
            const int ns = (int)eng.size();
            auto launch = [&] {
                check(launch_explicit_batched(rows_, reinterpret_cast<const ExplicitJob *>(d_blob_ + sec_exp_),
                                              d_starts_ + comp_.exp_starts_off, ns, comp_.total_tiles, stream_),
                      "explicit_batched");
            };
            launch();
            check(hipEventRecord(ev_[0], stream_), "event");
            for (int r = 0; r < repeats; ++r) launch();
            check(hipEventRecord(ev_[1], stream_), "event");
            check(hipEventSynchronize(ev_[1]), "event sync");
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "event");
            if (cand_ms[c] < 0.f || ms < cand_ms[c]) cand_ms[c] = ms;
        }
    int best_rows = cand[0];
    float best = cand_ms[0];
    for (int c = 1; c < 3; ++c)
        if (cand_ms[c] < best) { best = cand_ms[c]; best_rows = cand[c]; }
    rows_ = best_rows;
    comp_valid_ = false;
    if (tune_lines) tuneSharedLineShape(repeats);
    comp_valid_ = false;
    blob_used_[0] = blob_used_[1] = false;
    return rows_;
}

// The line solves run one launch per block-shape class.  Each member's own best shapes can
// leave several short launches whose tails idle the GPU, so the same lines under one shape
// shared by every member (a single launch) are timed too, and so are the chunk-size families:
// every line direction takes the fewest waves of one chunk size (4, 8 or 16 unknowns per
// thread) whose block holds it, so short lines keep small blocks in at most four launches.
// The faster plan is kept.
void Batch::tuneSharedLineShape(int repeats)
{
    std::vector<ShallowWaterEngine *> act;
    int nmax = 0;
    for (auto &kv : scen_)
        if (!paused_.count(kv.first)) {
            act.push_back(kv.second.get());
            nmax = std::max(nmax, kv.second->lineLength());
        }
    if (act.empty()) return;
    std::vector<std::pair<int, int>> own;
    for (auto *e : act) own.emplace_back(e->shapeX(), e->shapeY());
    auto time_solves = [&]() {
        std::vector<StepLaunch> plans;
        std::vector<ShallowWaterEngine *> eng;
        comp_valid_ = false;
        prepare(1, false, plans, eng);
        sync();
        char *h = h_blob_[0];
        fillTable(plans, eng, 1, h);
        check(hipMemcpy(d_blob_, h, stride_, hipMemcpyHostToDevice), "upload batch table");
        const SolveJob *sj = reinterpret_cast<const SolveJob *>(d_blob_ + sec_solve_);
        auto launch = [&] {
            for (int c = 0; c < kNumShapes; ++c)
                if (comp_.cls_count[c] > 0)
                    check(launch_solve_batched(c, sj + comp_.cls_first[c], d_starts_ + comp_.cls_starts_off[c],
                                               comp_.cls_count[c], comp_.cls_blocks[c], stream_), "solve_batched");
        };
        launch();
        check(hipEventRecord(ev_[0], stream_), "event");
        for (int r = 0; r < repeats; ++r) launch();
        check(hipEventRecord(ev_[1], stream_), "event");
        check(hipEventSynchronize(ev_[1]), "event sync");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "event");
        return ms;
    };
    // the smallest shape of chunk size C that holds n unknowns (-1: none)
    auto family_shape = [](int C, int n) {
        int best = -1;
        for (int sh = 0; sh < kNumShapes; ++sh)
            if (kShapeC[sh] == C && shape_capacity(sh) >= n && (best < 0 || kShapeW[sh] < kShapeW[best])) best = sh;
        return best;
    };
    // candidates: {shared shape or -1, chunk-size family or 0}; {-1, 0} is every member's own shapes
    std::vector<std::pair<int, int>> cands = {{-1, 0}};
    for (int C : {4, 8, 16})
        if (family_shape(C, nmax) >= 0) cands.emplace_back(-1, C);
    for (int sh = 0; sh < kNumShapes; ++sh)
        if (shape_capacity(sh) >= nmax && shape_capacity(sh) <= 4 * nmax) cands.emplace_back(sh, 0);
    auto apply = [&](const std::pair<int, int> &c) {
        for (size_t i = 0; i < act.size(); ++i) {
            if (c.first >= 0) act[i]->setLineShapes(c.first, c.first);
            else if (c.second > 0)
                act[i]->setLineShapes(family_shape(c.second, act[i]->lineLengthX()),
                                      family_shape(c.second, act[i]->lineLengthY()));
            else act[i]->setLineShapes(own[i].first, own[i].second);
        }
    };
    std::vector<float> ms(cands.size(), -1.f);
    for (int round = 0; round < 3; ++round)
        for (size_t c = 0; c < cands.size(); ++c) {
            apply(cands[c]);
            const float t = time_solves();
            if (ms[c] < 0.f || t < ms[c]) ms[c] = t;
        }
    size_t best = 0;
    for (size_t c = 1; c < cands.size(); ++c)
        if (ms[c] < ms[best]) best = c;
    apply(cands[best]);
}

void Batch::step(int nsteps)
{
    if (nsteps < 0) throw Error(CEL_EINVAL, "negative step count");
    int done = 0;
    if (graph_steps_ > 0)
        for (; done + graph_steps_ <= nsteps; done += graph_steps_) run(graph_steps_, true);
    // remainder: one table, direct launches (chunks bounded by the table's step capacity)
    while (done < nsteps) {
        const int m = std::min(nsteps - done, std::max(graph_steps_, 16));
        run(m, false);
        done += m;
    }
}

float Batch::timedSteps(int nsteps)
{
    check(hipEventRecord(ev_[0], stream_), "event");
    step(nsteps);
    check(hipEventRecord(ev_[1], stream_), "event");
    check(hipEventSynchronize(ev_[1]), "event sync");
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "event");
    return ms;
}

}  // namespace cel
