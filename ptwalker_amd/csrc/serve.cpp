// Serving loop (ptw_serve_replay): continuous batching of PPR requests over
// the resident graph.  Each launch carries the pilot walks of newly arrived
// requests, the main walks (and sampled 10x audits) of requests whose budgets
// are known, admitted in arrival order while their counter slices fit the
// HBM allowance.  Requests that arrive while a launch runs join the next one.
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

enum Kind : uint8_t { PILOT = 0, MAIN = 1, AUDIT = 2 };

struct Item {
    uint32_t req;
    Kind kind;
    uint64_t walkers;
};

// Slice bytes of one query under the batch planner's rule at the default
// dense threshold.
uint64_t slice_bytes(uint32_t V, uint64_t walkers, double tp)
{
    return ptw::batch_slice_sizes(V, walkers, tp, 100000).bytes();
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

    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };

    std::deque<uint32_t> pilot_q, ready_q;
    uint32_t next = 0, done = 0;
    ptw_serve_stats st{};
    std::vector<ptw_query> q;
    std::vector<Item> items;
    std::vector<uint32_t> tv;
    std::vector<uint64_t> tc, steps;
    std::vector<double> ent;
    std::vector<float> feats, pred;
    std::vector<uint32_t> gq, gv;
    std::vector<uint64_t> gout;

    while (done < n) {
        const double t = now_s();
        while (next < n && reqs[order[next]].arrival_s <= t) {
            const uint32_t r = order[next++];
            if (C.budget_mode == PTW_BUDGET_LEARNED) {
                pilot_q.push_back(r);
            } else {
                results[r].budget = fixed_budget(reqs[r].eps);
                ready_q.push_back(r);
            }
        }
        if (pilot_q.empty() && ready_q.empty()) {
            if (next < n) {
                const double wait = reqs[order[next]].arrival_s - now_s();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
            }
            continue;
        }
        // ---- compose one launch ----
        size_t f = 0, tot = 0;
        HIP_TRY(ctx, hipMemGetInfo(&f, &tot));
        const uint64_t allowance = C.hbm_limit ? C.hbm_limit : (uint64_t)(0.85 * (double)f);
        uint64_t used = 0;
        items.clear();
        while (!ready_q.empty() && items.size() + 2 <= C.max_batch) {
            const uint32_t r = ready_q.front();
            const bool aud = audit_pick(C.seed, reqs[r].id, C.audit_rate);
            const uint64_t b = results[r].budget;
            uint64_t need = slice_bytes(V, b, C.tp) + (aud ? slice_bytes(V, 10 * b, C.tp) : 0);
            if (used + need > allowance && !items.empty()) {
                ++st.admission_waits;
                break;
            }
            used += need;
            ready_q.pop_front();
            items.push_back(Item{r, MAIN, b});
            if (aud) items.push_back(Item{r, AUDIT, 10 * b});
            results[r].admit_s = t;
            results[r].audited = aud ? 1 : 0;
        }
        while (!pilot_q.empty() && items.size() < C.max_batch) {
            items.push_back(Item{pilot_q.front(), PILOT, kPilotWalkers});
            pilot_q.pop_front();
        }
        // ---- launch ----
        q.resize(items.size());
        for (size_t i = 0; i < items.size(); ++i) {
            const Item &it = items[i];
            const uint32_t id = reqs[it.req].id;
            const uint32_t stream = it.kind == MAIN ? id : it.kind == PILOT ? kPilotStream + id : kAuditStream + id;
            q[i] = ptw_query{reqs[it.req].source, stream, it.walkers, C.tp, PTW_QUERY_VISITS, 0};
            results[it.req].launches++;
        }
        ptw_batch_desc d{};
        d.queries = q.data();
        d.num_queries = (uint32_t)q.size();
        d.topk = K;
        d.seed = C.seed;
        d.semantics = PTW_SEM_REFERENCE;
        tv.assign(q.size() * K, 0);
        tc.assign(q.size() * K, 0);
        steps.assign(q.size(), 0);
        ent.assign(q.size(), 0.0);
        ptw_batch_out o{};
        o.topk_vertex = tv.data();
        o.topk_count = tc.data();
        o.steps = steps.data();
        o.entropy = ent.data();
        int rc = ptw_batch_ppr(ctx, &d, &o);
        if (rc) return rc;
        st.walk_steps += o.total_steps;
        st.launches++;
        const double t_done = now_s();
        // ---- pilots: features -> budgets ----
        feats.clear();
        std::vector<uint32_t> pilot_reqs;
        for (size_t i = 0; i < items.size(); ++i) {
            if (items[i].kind != PILOT) continue;
            const uint32_t r = items[i].req;
            float x[6];
            graph_features(ctx->host, reqs[r].source, x);
            double top10 = 0;
            for (uint32_t j = 0; j < std::min<uint32_t>(10, K); ++j) top10 += (double)tc[i * K + j];
            x[3] = (float)(top10 / ((double)kPilotWalkers + (double)steps[i]));
            x[4] = (float)ent[i];
            x[5] = std::log(reqs[r].eps);
            feats.insert(feats.end(), x, x + 6);
            pilot_reqs.push_back(r);
        }
        if (!pilot_reqs.empty()) {
            pred.assign(pilot_reqs.size(), 0.f);
            if ((rc = ptw_mlp_forward(mlp, feats.data(), (uint32_t)pilot_reqs.size(), pred.data()))) return rc;
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
        // ---- mains: answers; audits judge their main and top it up if short ----
        gq.clear();
        gv.clear();
        std::vector<std::pair<size_t, size_t>> pairs;  // (main item, audit item)
        for (size_t i = 0; i < items.size(); ++i) {
            if (items[i].kind != MAIN) continue;
            const uint32_t r = items[i].req;
            std::copy(tv.begin() + i * K, tv.begin() + (i + 1) * K, topk_vertex + (size_t)r * K);
            std::copy(tc.begin() + i * K, tc.begin() + (i + 1) * K, topk_count + (size_t)r * K);
            results[r].done_s = t_done;
            ++done;
            if (results[r].audited && i + 1 < items.size() && items[i + 1].kind == AUDIT) {
                pairs.emplace_back(i, i + 1);
                for (uint32_t j = 0; j < K; ++j) {
                    const uint32_t v = tv[(i + 1) * K + j];
                    gq.push_back((uint32_t)i);
                    gv.push_back(v == PTW_INVALID_VERTEX ? reqs[r].source : v);
                }
            }
        }
        if (!pairs.empty()) {
            gout.assign(gq.size(), 0);
            if ((rc = ptw_batch_gather(ctx, gq.size(), gq.data(), gv.data(), gout.data()))) return rc;
            for (size_t p = 0; p < pairs.size(); ++p) {
                const size_t mi = pairs[p].first, ai = pairs[p].second;
                const uint32_t r = items[mi].req;
                const double nm = (double)items[mi].walkers + (double)steps[mi];
                const double na = (double)items[ai].walkers + (double)steps[ai];
                double l1 = 0;
                for (uint32_t j = 0; j < K; ++j) {
                    if (tv[ai * K + j] == PTW_INVALID_VERTEX) continue;
                    l1 += std::fabs((double)gout[p * K + j] / nm - (double)tc[ai * K + j] / na);
                }
                st.audited++;
                if (l1 > (double)reqs[r].eps) {
                    results[r].topped_up = 1;
                    st.topped_up++;
                    std::copy(tv.begin() + ai * K, tv.begin() + (ai + 1) * K, topk_vertex + (size_t)r * K);
                    std::copy(tc.begin() + ai * K, tc.begin() + (ai + 1) * K, topk_count + (size_t)r * K);
                }
            }
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
