// pm_serve.cpp -- permeability tensor (three rotated solves) and the
// continuous-batching request loop, on top of the batch engine.
#include "../../include/poremaps_amd_serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

int pm_fail(int code, const char *fmt, ...);  // records the message for pm_last_error

namespace {

using clk = std::chrono::steady_clock;

// Rotation r of a (nz, ny, nx) x-fastest array so that physical axis
// flow_axis[r] becomes the array's z: r = 0 identity (flow along z),
// r = 1 transpose(A, (2, 0, 1)) (new z = x, y = z, x = y: flow along x),
// r = 2 transpose(A, (1, 2, 0)) (new z = y, y = x, x = z: flow along y).
// new_axis[r][a] = physical axis of the rotated array's axis a (0 = its x,
// 1 = its y, 2 = its z).
const int kNewAxis[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};

void rotate(const uint8_t *raw, int nx, int ny, int nz, int r, std::vector<uint8_t> &out, int dims[3]) {
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; a++) dims[a] = n[kNewAxis[r][a]];
    out.resize((size_t)nx * ny * nz);
    if (r == 0) {
        std::memcpy(out.data(), raw, out.size());
        return;
    }
    const int64_t stride[3] = {1, nx, (int64_t)nx * ny};  // physical strides of x, y, z
    const int64_t sa = stride[kNewAxis[r][0]], sb = stride[kNewAxis[r][1]], sc = stride[kNewAxis[r][2]];
    size_t o = 0;
    for (int c = 0; c < dims[2]; c++)
        for (int b = 0; b < dims[1]; b++) {
            const uint8_t *src = raw + c * sc + b * sb;
            for (int a = 0; a < dims[0]; a++) out[o++] = src[a * sa];
        }
}

// (wk13, wk23, wk33) of rotation r into k[3*i + j] (i velocity, j gradient).
void scatter(int r, const pm_log_row &row, double k[9]) {
    const int ax = kNewAxis[r][0], ay = kNewAxis[r][1], az = kNewAxis[r][2];
    k[3 * ax + az] = row.wk13;
    k[3 * ay + az] = row.wk23;
    k[3 * az + az] = row.wk33;
}

// The sub-domain (k13, k23, k33) of rotation r, scattered the same way.
void scatter_dom(int r, const pm_log_row &row, double k[9]) {
    const int ax = kNewAxis[r][0], ay = kNewAxis[r][1], az = kNewAxis[r][2];
    k[3 * ax + az] = row.k13;
    k[3 * ay + az] = row.k23;
    k[3 * az + az] = row.k33;
}

// A physical-axes box (x0 x1 y0 y1 z0 z1) in the axes of rotation r.
void rotate_box(const int box[6], int r, int out[6]) {
    for (int a = 0; a < 3; a++) {
        out[2 * a] = box[2 * kNewAxis[r][a]];
        out[2 * a + 1] = box[2 * kNewAxis[r][a] + 1];
    }
}

}  // namespace

extern "C" int pm_tensor(const pm_batch_params *p, const uint8_t *raw, int nx, int ny, int nz, int bc, double dx,
                         double phi, double k[9], pm_result res[3]) {
    if (!p || !raw || !k) return pm_fail(PM_E_ARG, "null argument");
    pm_batch *b = nullptr;
    int rc = pm_batch_create(p, &b);
    if (rc != PM_OK) return rc;
    int slot[3];
    std::vector<uint8_t> rot;
    for (int r = 0; r < 3 && rc == PM_OK; r++) {
        int d[3];
        rotate(raw, nx, ny, nz, r, rot, d);
        rc = pm_batch_add(b, rot.data(), d[0], d[1], d[2], bc, dx, phi, &slot[r]);
    }
    int running = 1;
    while (rc == PM_OK && running > 0) rc = pm_batch_step(b, &running);
    std::fill(k, k + 9, 0.0);
    for (int r = 0; r < 3 && rc == PM_OK; r++) {
        int st;
        pm_result rr;
        rc = pm_batch_status(b, slot[r], &st, &rr);
        if (rc == PM_OK) {
            scatter(r, rr.last, k);
            if (res) res[r] = rr;
        }
    }
    pm_batch_destroy(b);
    return rc;
}

extern "C" int pm_serve(const pm_serve_params *p, const pm_request *req, int n, pm_response *resp,
                        pm_serve_stats *stats) {
    if (!p || n < 0 || (!req && n) || (!resp && n)) return pm_fail(PM_E_ARG, "null argument");
    // Lanes: lane 0 takes every request, or with small_cells > 0 only those
    // of more voxels; lane 1 the requests of at most small_cells voxels.  Each
    // lane is a batch (one it_eval interval of all its solves per step); while
    // both hold solves, lane 1 steps for about as long as one lane-0 step
    // took, so a short request advances at the pace of the short ones and not
    // at that of the largest solve on the GPU.  Every solve runs the same
    // sweeps either way, so its result does not depend on its lane.
    struct Live {
        int req, rot;
    };
    struct Lane {
        pm_batch *b = nullptr;
        std::vector<Live> live;  // slot -> (request, rotation), req = -1 free
        int running = 0;
    };
    Lane lanes[2];
    const bool two = p->small_cells > 0;
    int rc = PM_OK;
    // one memory budget for both lanes, enforced below over their sum (each
    // lane's batch holds the whole budget as its own limit)
    pm_batch_params bp = p->batch;
    if (two && bp.budget_bytes <= 0) {
        // the default budget, as a batch computes it, fixed for both lanes
        pm_batch *probe = nullptr;
        rc = pm_batch_create(&bp, &probe);
        if (rc != PM_OK) return rc;
        int64_t u = 0;
        pm_batch_memory(probe, &u, &bp.budget_bytes);
        pm_batch_destroy(probe);
    }
    rc = pm_batch_create(&bp, &lanes[0].b);
    if (rc == PM_OK && two) rc = pm_batch_create(&bp, &lanes[1].b);
    if (rc != PM_OK) {
        for (Lane &L : lanes)
            if (L.b) pm_batch_destroy(L.b);
        return rc;
    }
    std::vector<int> left(n, 3);
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) {
        order[i] = i;
        std::memset(&resp[i], 0, sizeof(pm_response));
        resp[i].arrival_s = req[i].arrival_s;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return req[a].arrival_s < req[c].arrival_s; });
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    size_t next = 0;
    int done = 0, running = 0, max_running = 0, intervals = 0;
    double last_report = 0.0;
    int reported = 0;
    int64_t peak = 0;
    double admit_time = 0.0, step_time = 0.0;
    std::vector<uint8_t> rot;
    auto lane_of = [&](const pm_request &q) {
        return two && (int64_t)q.nx * q.ny * q.nz <= p->small_cells ? 1 : 0;
    };
    int rejected = 0;
    auto reject = [&](int id) {
        for (int r = 0; r < 3; r++) resp[id].state[r] = PM_SAMPLE_REJECTED;
        resp[id].admit_s = resp[id].done_s = now();
        left[id] = 0;
        rejected++;
        done++;
    };
    // the end of a launched interval of a lane: its finished solves evicted
    // and answered
    auto finish_lane = [&](Lane &L) -> int {
        int still = 0;
        const double t_step = now();
        int r = pm_batch_finish(L.b, &still);
        step_time += now() - t_step;
        intervals++;
        if (r != PM_OK) return r;
        bool progressed = still > 0;
        for (size_t s = 0; s < L.live.size(); s++) {
            if (L.live[s].req < 0) continue;
            int st;
            pm_result rr;
            if ((r = pm_batch_status(L.b, (int)s, &st, &rr)) != PM_OK) return r;
            if (st == PM_SAMPLE_RUNNING) continue;
            const int id = L.live[s].req, ro = L.live[s].rot;
            scatter(ro, rr.last, resp[id].k);
            scatter_dom(ro, rr.last, resp[id].k_dom);
            resp[id].sweeps[ro] = rr.sweeps;
            resp[id].state[ro] = st;
            pm_batch_remove(L.b, (int)s);
            L.live[s] = Live{-1, -1};
            L.running--;
            running--;
            progressed = true;
            if (--left[id] == 0) {
                resp[id].done_s = now();
                done++;
            }
        }
        if (!progressed) return pm_fail(PM_E_STATE, "admitted solves cannot progress");
        return PM_OK;
    };
    auto step_lane = [&](Lane &L) -> int {
        const double t_step = now();
        int r = pm_batch_launch(L.b);
        step_time += now() - t_step;
        return r == PM_OK ? finish_lane(L) : r;
    };
    while (rc == PM_OK && done < n) {
        // admission: first come, first served, while the three solves fit
        // their lane's budget
        while (next < order.size()) {
            const pm_request &q = req[order[next]];
            if (p->realtime && q.arrival_s > now()) break;
            Lane &L = lanes[lane_of(q)];
            int64_t used = 0, budget = 0, used1 = 0, need = 0;
            pm_batch_memory(lanes[0].b, &used, &budget);
            if (lanes[1].b) pm_batch_memory(lanes[1].b, &used1, nullptr);
            used += used1;
            for (int r = 0; r < 3; r++) {
                int d[3];
                const int nn[3] = {q.nx, q.ny, q.nz};
                for (int a = 0; a < 3; a++) d[a] = nn[kNewAxis[r][a]];
                need += pm_batch_sample_bytes(&p->batch, d[0], d[1], d[2]);
            }
            const int id = order[next];
            if (need > budget) {
                // can never fit, even alone: answered as rejected
                reject(id);
                next++;
                continue;
            }
            if (used + need > budget && running > 0) break;  // back-pressure
            if (p->max_solves > 0 && running > 0 && running + 3 > p->max_solves) break;
            resp[id].admit_s = now();
            const double t_admit = now();
            for (int r = 0; r < 3 && rc == PM_OK; r++) {
                int d[3], slot = -1;
                rotate(q.raw, q.nx, q.ny, q.nz, r, rot, d);
                rc = pm_batch_add(L.b, rot.data(), d[0], d[1], d[2], q.boundary_method, q.voxel_size, q.porosity,
                                  &slot);
                if (rc == PM_OK && q.eps > 0.0) rc = pm_batch_set_eps(L.b, slot, q.eps);
                if (rc == PM_OK) {
                    int box[6];
                    rotate_box(q.dom_interest, r, box);
                    rc = pm_batch_set_dom_interest(L.b, slot, box);
                }
                if (rc != PM_OK) {
                    if (slot >= 0) pm_batch_remove(L.b, slot);
                    break;
                }
                if ((int)L.live.size() <= slot) L.live.resize(slot + 1, Live{-1, -1});
                L.live[slot] = Live{id, r};
                L.running++;
                running++;
            }
            if (rc == PM_E_MEMORY || rc == PM_E_ARG || rc == PM_E_GEOMETRY) {
                // this request alone is refused: its admitted rotations leave
                // the batch, the others keep being served
                for (size_t s = 0; s < L.live.size(); s++)
                    if (L.live[s].req == id) {
                        pm_batch_remove(L.b, (int)s);
                        L.live[s] = Live{-1, -1};
                        L.running--;
                        running--;
                    }
                reject(id);
                rc = PM_OK;
            }
            next++;
            admit_time += now() - t_admit;
            int64_t u2 = 0, u3 = 0;
            pm_batch_memory(lanes[0].b, &u2, nullptr);
            if (lanes[1].b) pm_batch_memory(lanes[1].b, &u3, nullptr);
            peak = std::max(peak, u2 + u3);
        }
        if (rc != PM_OK) break;
        max_running = std::max(max_running, running);
        if (running == 0) {
            if (next < order.size()) {
                const double wait = req[order[next]].arrival_s - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            }
            continue;
        }
        // time slices: one lane-0 interval, then lane 1 for about as long.
        // (Stepping lane 1 on its stream while a lane-0 interval runs leaves it
        // only the CUs the large interval frees as its workgroups retire: on
        // the serving workload the same throughput with p50 11.4 s instead of
        // 1.4 s.)
        double t_large = 0.0;
        if (lanes[0].running > 0) {
            const double ts = now();
            rc = step_lane(lanes[0]);
            t_large = now() - ts;
        }
        if (rc == PM_OK && lanes[1].running > 0) {
            const double ts = now();
            int k = 0;
            do {
                rc = step_lane(lanes[1]);
                k++;
            } while (rc == PM_OK && lanes[1].running > 0 && lanes[0].running > 0 && now() - ts < t_large &&
                     k < 64);
        }
        if (rc == PM_OK && p->progress && (done > reported || now() - last_report >= 1.0)) {
            reported = done;
            last_report = now();
            if (p->progress(done, running, last_report, p->user)) rc = pm_fail(PM_E_ABORTED, "stopped by the caller");
        }
    }
    for (Lane &L : lanes)
        if (L.b) pm_batch_destroy(L.b);
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        std::vector<double> lat;
        double first = 1e300, last = 0.0;
        for (int i = 0; i < n; i++) {
            if (resp[i].done_s <= 0.0 || resp[i].state[0] == PM_SAMPLE_REJECTED) continue;
            lat.push_back(resp[i].done_s - resp[i].arrival_s);
            first = std::min(first, resp[i].arrival_s);
            last = std::max(last, resp[i].done_s);
        }
        std::sort(lat.begin(), lat.end());
        stats->completed = (int)lat.size();
        if (!lat.empty()) {
            auto pct = [&](double f) {
                const double x = f * (lat.size() - 1);
                const size_t lo = (size_t)std::floor(x), hi = std::min(lat.size() - 1, lo + 1);
                return lat[lo] + (lat[hi] - lat[lo]) * (x - lo);
            };
            stats->latency_p50_s = pct(0.50);
            stats->latency_p99_s = pct(0.99);
            double s = 0;
            for (double v : lat) s += v;
            stats->latency_mean_s = s / lat.size();
            stats->latency_max_s = lat.back();
            stats->wall_s = last - std::min(first, last);
            stats->solves_per_hour = 3.0 * lat.size() / std::max(stats->wall_s, 1e-9) * 3600.0;
        }
        stats->max_running = max_running;
        stats->peak_bytes = peak;
        stats->intervals = intervals;
        stats->admit_s = admit_time;
        stats->step_s = step_time;
        stats->rejected = rejected;
    }
    return rc;
}
