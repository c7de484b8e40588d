// Serving loop (ptw_serve_replay): continuous batching of PPR requests over
// the resident graph.  Two lanes (each with its own stream, counter pools and
// half of the CUs for its persistent kernels) carry one batch each; a lane
// that is idle takes the next batch at once, so requests that arrive while a
// batch runs start on the device beside it.  A batch carries the main walks
// (and sampled 10x audits) of requests whose budgets are known, admitted in
// arrival order while their counter slices fit the lane's HBM allowance, and
// the pilot walks of newly arrived requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ptw.h"
#include "ctx.hpp"
#include "philox.hpp"

namespace {

using ptw::fail;

constexpr uint32_t kPilotWalkers = 1000;
constexpr uint32_t kPilotStream = 1u << 30;
constexpr uint32_t kAuditStream = 1u << 31;
// Walkers a batch takes at most (its first request excepted): about 60 ms of
// walking on the whole device, so a large backlog is served as a run of
// batches that alternate between the lanes rather than as one long batch.
constexpr double kBatchWalkers = 4e8;

enum Kind : uint8_t { PILOT = 0, MAIN = 1, AUDIT = 2 };

struct Item {
    uint32_t req;
    Kind kind;
    uint64_t walkers;
};

// HBM of one query under the batch planner's rule at the default dense
// threshold (logged dense queries when the context logs them).
uint64_t slice_bytes(const ptw_ctx *ctx, uint64_t walkers, double tp)
{
    const uint32_t V = ctx->dev.V;
    const bool logged = ctx->shape.visit_log && ptw::batch_log_fits(V);
    const uint64_t thr = logged ? ptw::kDenseThresholdLogged : ptw::kDenseThresholdSlices;
    const uint32_t hubs =
        logged ? std::min<uint32_t>(ptw::batch_log_hubs(ctx->shape.lds_hubs, ((V - 1) >> ptw::kBatchLogShift) + 1,
                                                         ptw::batch_cache_slots(ctx->shape)),
                                    ptw::dense_pitch(V))
               : 0u;
    return ptw::batch_query_bytes(V, walkers, tp, false, thr, logged, hubs);
}

void graph_features(const ptw::HostCsr &g, uint32_t s, float *f)
{
    const uint32_t b = g.row_ptr[s], e = g.row_ptr[s + 1];
    const double d = e - b;
    double two_hop = 0.0;
    for (uint32_t k = b; k < e; ++k) two_hop += g.row_ptr[g.col[k] + 1] - g.row_ptr[g.col[k]];
    f[0] = (float)std::log(std::max(d, 1.0));
    f[1] = (float)std::log(std::max(two_hop / std::max(d, 1.0), 1.0));
    f[2] = (float)std::log(std::max(two_hop, 1.0));
}

// The launch a lane carries: its items and the query table they became.
struct InFlight {
    bool busy = false;
    std::vector<Item> items;
    std::vector<ptw_query> q;
    ptw_batch_desc d{};
};

// Streams (default priority), events and result staging of the serving lanes,
// made on the first serving call and kept with the context.
int ensure_lanes(ptw_ctx *ctx)
{
    for (auto &l : ctx->serve) {
        if (l.stream) continue;
        HIP_TRY(ctx, hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));
        HIP_TRY(ctx, hipEventCreate(&l.ev0));
        HIP_TRY(ctx, hipEventCreate(&l.ev1));
        HIP_TRY(ctx, hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
    }
    return PTW_OK;
}

// Page-locked staging for m queries with k top-k entries each.
int grow_staging(ptw_ctx *ctx, ptw::ServeLane &l, uint64_t m, uint64_t k)
{
    m = std::max<uint64_t>(m, 1);
    if (m > l.h_q) {
        if (l.h_steps) HIP_TRY(ctx, hipHostFree(l.h_steps));
        if (l.h_over) HIP_TRY(ctx, hipHostFree(l.h_over));
        if (l.h_ent) HIP_TRY(ctx, hipHostFree(l.h_ent));
        l.h_steps = nullptr;
        l.h_over = nullptr;
        l.h_ent = nullptr;
        l.h_q = 0;
        HIP_TRY(ctx, hipHostMalloc((void **)&l.h_steps, m * 8, hipHostMallocDefault));
        HIP_TRY(ctx, hipHostMalloc((void **)&l.h_over, m * 4, hipHostMallocDefault));
        HIP_TRY(ctx, hipHostMalloc((void **)&l.h_ent, m * 8, hipHostMallocDefault));
        l.h_q = m;
    }
    const uint64_t mk = std::max<uint64_t>(m * k, 1);
    if (mk > l.h_k) {
        if (l.h_tv) HIP_TRY(ctx, hipHostFree(l.h_tv));
        if (l.h_tc) HIP_TRY(ctx, hipHostFree(l.h_tc));
        l.h_tv = nullptr;
        l.h_tc = nullptr;
        l.h_k = 0;
        HIP_TRY(ctx, hipHostMalloc((void **)&l.h_tv, mk * 4, hipHostMallocDefault));
        HIP_TRY(ctx, hipHostMalloc((void **)&l.h_tc, mk * 8, hipHostMallocDefault));
        l.h_k = mk;
    }
    return PTW_OK;
}

bool audit_pick(uint64_t seed, uint32_t req, double rate)
{
    if (rate <= 0.0) return false;
    const ptw::u32x4 r = ptw::philox4x32_10(req, 0, 0, 0x41554449u, (uint32_t)seed, (uint32_t)(seed >> 32));
    return (double)r.x < rate * 4294967296.0;
}

}  // namespace

extern "C" {

int ptw_serve_replay(ptw_ctx *ctx, ptw_mlp *mlp, const ptw_serve_config *cfg, const ptw_request *reqs, uint32_t n,
                     ptw_request_result *results, uint32_t *topk_vertex, uint64_t *topk_count,
                     ptw_serve_stats *stats)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!cfg || (n && (!reqs || !results))) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: serving needs a HIP device");
    const ptw_serve_config &C = *cfg;
    if (C.budget_mode != PTW_BUDGET_FIXED && C.budget_mode != PTW_BUDGET_LEARNED)
        return fail(ctx, PTW_ERR_INVALID_ARG, "unknown budget mode");
    if (C.budget_mode == PTW_BUDGET_LEARNED && !mlp) return fail(ctx, PTW_ERR_INVALID_ARG, "learned mode needs a model");
    if (!(C.tp > 0.0 && C.tp <= 1.0)) return fail(ctx, PTW_ERR_INVALID_ARG, "tp must be in (0, 1]");
    if (C.topk == 0 || C.topk > 1024 || 2 * C.topk > (uint32_t)ctx->shape.topk_tile)
        return fail(ctx, PTW_ERR_INVALID_ARG, "topk must be in [1, 1024] and <= topk_tile / 2");
    if (C.topk && n && (!topk_vertex || !topk_count)) return fail(ctx, PTW_ERR_INVALID_ARG, "top-k outputs required");
    if (C.max_batch < 2) return fail(ctx, PTW_ERR_INVALID_ARG, "max_batch must be >= 2");
    if (C.w_min == 0 || C.w_max < C.w_min) return fail(ctx, PTW_ERR_INVALID_ARG, "bad budget clip");
    if (C.budget_mode == PTW_BUDGET_LEARNED && C.topk < 10)
        return fail(ctx, PTW_ERR_INVALID_ARG, "learned mode needs topk >= 10 (pilot top-10 mass)");
    if (!(C.audit_rate >= 0.0 && C.audit_rate <= 1.0)) return fail(ctx, PTW_ERR_INVALID_ARG, "audit_rate must be in [0, 1]");
    // the largest walk (the largest budget, or its 10x audit when audits are on)
    // must fit the u32 counter slices
    const double largest = (C.audit_rate > 0.0 ? 10.0 : 1.0) * (double)C.w_max;
    if (largest * (1.0 + 1.0 / C.tp) * 1.5 > 4294967296.0)
        return fail(ctx, PTW_ERR_INVALID_ARG, "w_max too large for u32 counter slices at this tp");
    const uint32_t V = ctx->dev.V, K = C.topk;
    for (uint32_t i = 0; i < n; ++i) {
        if (reqs[i].source >= V) return fail(ctx, PTW_ERR_INVALID_ARG, "request source out of range");
        if (reqs[i].id >= kPilotStream) return fail(ctx, PTW_ERR_INVALID_ARG, "request id must be < 2^30");
        if (!(reqs[i].eps > 0.f)) return fail(ctx, PTW_ERR_INVALID_ARG, "request eps must be > 0");
        if (!(reqs[i].arrival_s >= 0.0)) return fail(ctx, PTW_ERR_INVALID_ARG, "arrival_s must be >= 0");
    }
    for (uint32_t i = 0; i < n; ++i) results[i] = ptw_request_result{};

    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    auto fixed_budget = [&](float eps) {
        double w = std::ceil(C.fixed_c2 / ((double)eps * (double)eps));
        return (uint64_t)std::min<double>(std::max<double>(w, (double)C.w_min), (double)C.w_max);
    };

    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (int rc = ensure_lanes(ctx)) return rc;
    // a lane's batch that takes all waiting work fills half the CUs with its
    // persistent kernels, so a batch launched on the idle lane runs beside it
    const int lane_cus = std::max(1, ctx->num_cus / ptw::kServeLanes);
    size_t f0 = 0, tot0 = 0;
    HIP_TRY(ctx, hipMemGetInfo(&f0, &tot0));
    const uint64_t allowance = (C.hbm_limit ? C.hbm_limit : (uint64_t)(0.85 * (double)f0)) / ptw::kServeLanes;

    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };

    std::deque<uint32_t> pilot_q, ready_q;
    uint32_t next = 0, done = 0;
    ptw_serve_stats st{};
    InFlight fl[ptw::kServeLanes];
    std::vector<float> feats, pred;
    std::vector<uint32_t> gq, gv;
    std::vector<uint64_t> gout;

    auto item_walkers = [&](uint32_t r) {
        return (double)results[r].budget * (audit_pick(C.seed, reqs[r].id, C.audit_rate) ? 11.0 : 1.0);
    };
    // Composes one batch for an idle lane and queues it (no wait): the mains
    // (and sampled audits) of ready requests while their slices fit the lane's
    // HBM allowance, then the pilots of new arrivals.
    //  - A batch that leaves work waiting (a backlog: latency is set by the
    //    queue) fills every CU and queues behind the other lane's batch.
    //  - With `split` (both lanes idle, the waiting work fits one batch), the
    //    batch takes about half of it and the other lane the rest, each on
    //    half of the CUs, side by side.
    //  - Otherwise it takes all waiting work on half of the CUs, so a request
    //    that arrives meanwhile starts at once on the other lane.
    auto launch = [&](int li, double t, bool split) -> int {
        ptw::ServeLane &SL = ctx->serve[li];
        InFlight &F = fl[li];
        F.items.clear();
        double waiting = 0;
        if (split)
            for (uint32_t r : ready_q) waiting += item_walkers(r);
        const size_t pilots = split ? (pilot_q.size() + 1) / 2 : pilot_q.size();
        uint64_t used = 0;
        double walkers = 0;
        while (!ready_q.empty() && F.items.size() + 2 <= C.max_batch) {
            const uint32_t r = ready_q.front();
            const bool aud = audit_pick(C.seed, reqs[r].id, C.audit_rate);
            const uint64_t b = results[r].budget;
            const double w = (double)b * (aud ? 11.0 : 1.0);
            if (!F.items.empty() && (walkers + w > kBatchWalkers || (split && 2 * walkers >= waiting))) break;
            walkers += w;
            uint64_t need = slice_bytes(ctx, b, C.tp) + (aud ? slice_bytes(ctx, 10 * b, C.tp) : 0);
            if (used + need > allowance && !F.items.empty()) {
                ++st.admission_waits;
                break;
            }
            used += need;
            ready_q.pop_front();
            F.items.push_back(Item{r, MAIN, b});
            if (aud) F.items.push_back(Item{r, AUDIT, 10 * b});
            results[r].admit_s = t;
            results[r].audited = aud ? 1 : 0;
        }
        for (size_t k = 0; k < pilots && !pilot_q.empty() && F.items.size() < C.max_batch; ++k) {
            F.items.push_back(Item{pilot_q.front(), PILOT, kPilotWalkers});
            pilot_q.pop_front();
        }
        const bool backlog = !split && (!ready_q.empty() || !pilot_q.empty());
        const int cus = backlog ? ctx->num_cus : lane_cus;
        const size_t m = F.items.size();
        F.q.resize(m);
        for (size_t i = 0; i < m; ++i) {
            const Item &it = F.items[i];
            const uint32_t id = reqs[it.req].id;
            const uint32_t stream = it.kind == MAIN ? id : it.kind == PILOT ? kPilotStream + id : kAuditStream + id;
            F.q[i] = ptw_query{reqs[it.req].source, stream, it.walkers, C.tp, PTW_QUERY_VISITS, 0};
            results[it.req].launches++;
        }
        F.d = ptw_batch_desc{};
        F.d.queries = F.q.data();
        F.d.num_queries = (uint32_t)m;
        F.d.topk = K;
        F.d.seed = C.seed;
        F.d.semantics = PTW_SEM_REFERENCE;
        if (int rc = grow_staging(ctx, SL, m, K)) return rc;
        const ptw::LaneRef L{SL.b, SL.stream, SL.ev0, SL.ev1, cus};
        if (int rc = ptw::batch_enqueue(ctx, L, F.d)) return rc;
        if (int rc = ptw::batch_enqueue_results(ctx, L, (uint32_t)m, K, SL.h_steps, SL.h_over, SL.h_ent, SL.h_tv,
                                                SL.h_tc))
            return rc;
        HIP_TRY(ctx, hipEventRecord(SL.done, SL.stream));
        F.busy = true;
        st.launches++;
        return PTW_OK;
    };

    // A lane's batch has completed: budgets from its pilots, answers from its
    // mains, and each audit's verdict on its main.
    auto complete = [&](int li, double t_done) -> int {
        ptw::ServeLane &SL = ctx->serve[li];
        InFlight &F = fl[li];
        F.busy = false;
        const size_t m = F.items.size();
        if (int rc = ptw::batch_check(ctx, F.d, SL.h_steps, SL.h_over)) return rc;
        for (size_t i = 0; i < m; ++i) st.walk_steps += SL.h_steps[i];
        feats.clear();
        std::vector<uint32_t> pilot_reqs;
        for (size_t i = 0; i < m; ++i) {
            if (F.items[i].kind != PILOT) continue;
            const uint32_t r = F.items[i].req;
            float x[6];
            graph_features(ctx->host, reqs[r].source, x);
            double top10 = 0;
            for (uint32_t j = 0; j < std::min<uint32_t>(10, K); ++j) top10 += (double)SL.h_tc[i * K + j];
            x[3] = (float)(top10 / ((double)kPilotWalkers + (double)SL.h_steps[i]));
            x[4] = (float)SL.h_ent[i];
            x[5] = std::log(reqs[r].eps);
            feats.insert(feats.end(), x, x + 6);
            pilot_reqs.push_back(r);
        }
        if (!pilot_reqs.empty()) {
            pred.assign(pilot_reqs.size(), 0.f);
            if (int rc = ptw_mlp_forward(mlp, feats.data(), (uint32_t)pilot_reqs.size(), pred.data())) return rc;
            for (size_t j = 0; j < pilot_reqs.size(); ++j) {
                const uint32_t r = pilot_reqs[j];
                bool ood = false;
                for (int k = 0; k < 6; ++k) ood |= feats[j * 6 + k] < C.feat_lo[k] || feats[j * 6 + k] > C.feat_hi[k];
                uint64_t b;
                if (ood) {
                    b = fixed_budget(reqs[r].eps);
                    results[r].fallback = 1;
                    st.fallbacks++;
                } else {
                    const double w = std::ceil(C.margin * std::exp((double)pred[j]));
                    b = (uint64_t)std::min<double>(std::max<double>(w, (double)C.w_min), (double)C.w_max);
                }
                results[r].budget = b;
                ready_q.push_back(r);
            }
        }
        gq.clear();
        gv.clear();
        std::vector<std::pair<size_t, size_t>> pairs;  // (main item, audit item)
        for (size_t i = 0; i < m; ++i) {
            if (F.items[i].kind != MAIN) continue;
            const uint32_t r = F.items[i].req;
            std::copy(SL.h_tv + i * K, SL.h_tv + (i + 1) * K, topk_vertex + (size_t)r * K);
            std::copy(SL.h_tc + i * K, SL.h_tc + (i + 1) * K, topk_count + (size_t)r * K);
            results[r].done_s = t_done;
            ++done;
            if (results[r].audited && i + 1 < m && F.items[i + 1].kind == AUDIT) {
                pairs.emplace_back(i, i + 1);
                for (uint32_t j = 0; j < K; ++j) {
                    const uint32_t v = SL.h_tv[(i + 1) * K + j];
                    gq.push_back((uint32_t)i);
                    gv.push_back(v == PTW_INVALID_VERTEX ? reqs[r].source : v);
                }
            }
        }
        if (!pairs.empty()) {
            gout.assign(gq.size(), 0);
            const ptw::LaneRef L{SL.b, SL.stream, SL.ev0, SL.ev1, lane_cus};
            if (int rc = ptw::batch_gather_lane(ctx, L, gq.size(), gq.data(), gv.data(), gout.data())) return rc;
            for (size_t p = 0; p < pairs.size(); ++p) {
                const size_t mi = pairs[p].first, ai = pairs[p].second;
                const uint32_t r = F.items[mi].req;
                const double nm = (double)F.items[mi].walkers + (double)SL.h_steps[mi];
                const double na = (double)F.items[ai].walkers + (double)SL.h_steps[ai];
                double l1 = 0;
                for (uint32_t j = 0; j < K; ++j) {
                    if (SL.h_tv[ai * K + j] == PTW_INVALID_VERTEX) continue;
                    l1 += std::fabs((double)gout[p * K + j] / nm - (double)SL.h_tc[ai * K + j] / na);
                }
                st.audited++;
                if (l1 > (double)reqs[r].eps) {
                    results[r].topped_up = 1;
                    st.topped_up++;
                    std::copy(SL.h_tv + ai * K, SL.h_tv + (ai + 1) * K, topk_vertex + (size_t)r * K);
                    std::copy(SL.h_tc + ai * K, SL.h_tc + (ai + 1) * K, topk_count + (size_t)r * K);
                }
            }
        }
        return PTW_OK;
    };

    while (done < n) {
        bool progressed = false;
        // completions first: they free lanes and make budgets known
        for (int li = 0; li < ptw::kServeLanes; ++li) {
            if (!fl[li].busy) continue;
            const hipError_t e = hipEventQuery(ctx->serve[li].done);
            if (e == hipErrorNotReady) continue;
            HIP_TRY(ctx, e);
            if (int rc = complete(li, now_s())) return rc;
            progressed = true;
        }
        const double t = now_s();
        while (next < n && reqs[order[next]].arrival_s <= t) {
            const uint32_t r = order[next++];
            if (C.budget_mode == PTW_BUDGET_LEARNED) {
                pilot_q.push_back(r);
            } else {
                results[r].budget = fixed_budget(reqs[r].eps);
                ready_q.push_back(r);
            }
            progressed = true;
        }
        for (int li = 0; li < ptw::kServeLanes && (!pilot_q.empty() || !ready_q.empty()); ++li) {
            if (fl[li].busy) continue;
            bool other_idle = false;
            for (int lj = li + 1; lj < ptw::kServeLanes; ++lj) other_idle |= !fl[lj].busy;
            bool split = false;
            if (other_idle && ready_q.size() + pilot_q.size() >= 2 && ready_q.size() + 2 * pilot_q.size() <= C.max_batch) {
                double waiting = 0;
                for (uint32_t r : ready_q) waiting += item_walkers(r);
                split = waiting <= kBatchWalkers;
            }
            if (int rc = launch(li, t, split)) return rc;
            progressed = true;
        }
        if (!progressed) {
            // nothing to do until a lane completes or the next request arrives
            double wait = 50e-6;
            if (next < n) wait = std::min(wait, std::max(0.0, reqs[order[next]].arrival_s - now_s()));
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        }
    }
    // ---- statistics ----
    std::vector<double> lat(n);
    double last = 0, first = n ? reqs[order[0]].arrival_s : 0;
    for (uint32_t i = 0; i < n; ++i) {
        lat[i] = results[i].done_s - reqs[i].arrival_s;
        last = std::max(last, results[i].done_s);
    }
    std::sort(lat.begin(), lat.end());
    if (stats) {
        st.makespan_s = last - first;
        st.queries_per_s = st.makespan_s > 0 ? n / st.makespan_s : 0;
        if (n) {
            st.p50_ms = 1e3 * lat[(size_t)(0.50 * (n - 1))];
            st.p99_ms = 1e3 * lat[(size_t)(0.99 * (n - 1))];
            st.mean_ms = 1e3 * std::accumulate(lat.begin(), lat.end(), 0.0) / n;
        }
        *stats = st;
    }
    return PTW_OK;
}

}  // extern "C"
