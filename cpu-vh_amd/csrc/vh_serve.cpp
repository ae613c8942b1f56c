// Serving loop of the batch engine (include/cpuvh_amd.h, vh_serve_run): an
// open-loop request stream admitted by HBM footprint into one vh_batch,
// stepped in rounds of `coarse` steps until each request meets the
// reference's stop rule (rhic/src/HydroPlugin.cpp:266-281: the third
// freezeout check with no cell above e_F; the count of such checks is never
// reset).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/cpuvh_amd.h"
#include "vh_host_eos.hpp"

namespace {

struct Rng {  // SplitMix64
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    int uniform_int(int lo, int hi) { return lo + (int)(uniform() * (hi - lo + 1)); }
};

struct Payload {
    std::vector<double> e, p;
};

// Hash of an event's state bits (the vh_serve_request digest): fields in
// vh_download's numbering, each weighted by position so that order counts.
uint64_t state_digest(vh_batch *b, int id, int nc, int64_t cells, std::vector<double> &buf) {
    uint64_t h = 14695981039346656037ull;
    std::vector<int> fields;
    for (int f = 0; f < nc; ++f) fields.push_back(f);
    for (int f = 15; f <= 20; ++f) fields.push_back(f);
    for (int f : fields) {
        if (vh_batch_download(b, id, f, buf.data()) != VH_OK) return 0;
        uint64_t d = 0;
        for (int64_t i = 0; i < cells; ++i) {
            uint64_t bits;
            std::memcpy(&bits, &buf[i], sizeof bits);
            d += bits * (2 * (uint64_t)i + 1);
        }
        h = h * 1099511628211ull + d;
    }
    return h;
}

// A fixed set of host threads running parts of one job at a time: run(n, f)
// calls f(0) .. f(n - 1), the caller taking parts too, and returns when all
// are done.
class HostPool {
  public:
    explicit HostPool(int threads) {
        for (int k = 0; k < threads; ++k) th_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            quit_ = true;
        }
        go_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(int parts, const std::function<void(int)> &f) {
        if (th_.empty() || parts < 2) {
            for (int k = 0; k < parts; ++k) f(k);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &f;
            parts_ = parts;
            next_ = 0;
            left_ = parts;
            ++gen_;
        }
        go_.notify_all();
        work();
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [this] { return left_ == 0; });
        job_ = nullptr;
        parts_ = 0;
    }

  private:
    // Parts are taken under the lock, so a thread still leaving the last job
    // can only take parts of the current one.
    void work() {
        for (;;) {
            const std::function<void(int)> *job;
            int k;
            {
                std::lock_guard<std::mutex> g(m_);
                if (!job_ || next_ >= parts_) return;
                k = next_++;
                job = job_;
            }
            (*job)(k);
            std::lock_guard<std::mutex> g(m_);
            if (--left_ == 0) done_.notify_all();
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                go_.wait(g, [&] { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
            }
            work();
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable go_, done_;
    const std::function<void(int)> *job_ = nullptr;
    int parts_ = 0, left_ = 0, next_ = 0;
    uint64_t gen_ = 0;
    bool quit_ = false;
};

// The surfaces of the requests in batch slots sel at one check, in one pass
// (vh_batch_fo_surfaces): each slot's element counts ne, elements el,
// variables var and element count tot.  A slot whose arrays were too small
// gets larger ones and is built again in a second pass.
vh_status surface_pass(vh_batch *b, const std::vector<int> &sel, const vh_serve_config &c,
                       const std::vector<int64_t> &nflag, std::vector<std::vector<int>> &ne,
                       std::vector<std::vector<double>> &el, std::vector<std::vector<double>> &var,
                       std::vector<int64_t> &tot) {
    std::vector<int> ids(sel), again;
    for (int pass = 0; pass < 2 && !ids.empty(); ++pass) {
        const size_t n = ids.size();
        std::vector<double> widths(4 * n);
        std::vector<int *> pn(n);
        std::vector<double *> pe(n), pv(n);
        std::vector<int64_t> cap(n), got(n);
        for (size_t k = 0; k < n; ++k) {
            const int id = ids[k];
            widths[4 * k] = c.dt * c.coarse;
            widths[4 * k + 1] = widths[4 * k + 2] = widths[4 * k + 3] = c.dx;
            ne[id].resize((size_t)nflag[id]);
            if (pass == 0 && var[id].size() < (size_t)nflag[id] * 20) {  // about one element a cube
                el[id].resize((size_t)nflag[id] * 16);
                var[id].resize((size_t)nflag[id] * 20);
            }
            pn[k] = ne[id].data();
            pe[k] = el[id].data();
            pv[k] = var[id].data();
            cap[k] = (int64_t)(var[id].size() / 10);  // el holds 8 doubles an element
        }
        vh_status s = vh_batch_fo_surfaces(b, (int)n, ids.data(), widths.data(), pn.data(), pe.data(), pv.data(),
                                           cap.data(), got.data());
        if (s != VH_OK && s != VH_ERR_CAPACITY) return s;
        again.clear();
        for (size_t k = 0; k < n; ++k) {
            const int id = ids[k];
            tot[id] = got[k];
            if (got[k] > cap[k]) {
                el[id].resize((size_t)got[k] * 8);
                var[id].resize((size_t)got[k] * 10);
                again.push_back(id);
            }
        }
        ids.swap(again);
    }
    return ids.empty() ? VH_OK : VH_ERR_STATE;
}

// One request's surface at a check as surface.dat's 20 columns per element
// (FreezeOut.cpp:200-361): cubes in the reference's order (x outer, eta
// inner), positions from the cube's origin, dsigma = t * normal, T and P from
// the host EoS at the interpolated e; steps is the request's step count at
// the check; ids, ne, el, var and total as the flag and surface_pass left
// them.
void surface_rows(const vh_serve_request &r, const vh_serve_config &c, const vh::HostEos &heos, int steps,
                  const std::vector<int64_t> &ids, int64_t nf, const std::vector<int> &ne,
                  const std::vector<double> &el, const std::vector<double> &var, int64_t total,
                  std::vector<double> &rows, std::vector<int64_t> &order, std::vector<int64_t> &bucket,
                  HostPool &pool) {
    const bool d4 = r.nz > 1;
    const int dim = d4 ? 4 : 3, per = 2 * dim;
    const double t = c.t0 + steps * c.dt, cell_tau = t + c.dt;
    // The cubes come flagged in increasing id (x fastest); the reference lists
    // them x outermost, then y, then eta.  A stable counting sort on
    // ix * cy + iy gives that order: within one (ix, iy) the ids increase
    // with eta.
    const int64_t cx = r.nx - 1, cy = r.ny - 1, cz = d4 ? r.nz - 1 : 1;
    (void)cz;
    bucket.assign((size_t)(cx * cy + 1), 0);
    for (int64_t m = 0; m < nf; ++m) {
        const int64_t q = ids[m];
        ++bucket[(size_t)((q % cx) * cy + (q / cx) % cy) + 1];
    }
    for (size_t k = 1; k < bucket.size(); ++k) bucket[k] += bucket[k - 1];
    order.resize((size_t)nf * 2);  // [cube in the reference's order | its first row]
    for (int64_t m = 0; m < nf; ++m) {
        const int64_t q = ids[m];
        order[(size_t)bucket[(size_t)((q % cx) * cy + (q / cx) % cy)]++] = m;
    }
    std::vector<int64_t> off(nf + 1, 0);
    for (int64_t m = 0; m < nf; ++m) off[m + 1] = off[m] + ne[m];
    int64_t *first = order.data() + nf;
    for (int64_t k = 0, at = 0; k < nf; ++k) {
        first[k] = at;
        at += ne[order[k]];
    }
    rows.resize((size_t)total * 20);
    // the rows of consecutive runs of cubes, on the pool's threads
    const int parts = (int)std::min<int64_t>(nf, 64);
    pool.run(parts, [&](int part) {
        const int64_t k0 = nf * part / parts, k1 = nf * (part + 1) / parts;
        for (int64_t k = k0; k < k1; ++k) {
            const int64_t m = order[k], q = ids[m];
            const int ix = (int)(q % cx), iy = (int)((q / cx) % cy), iz = d4 ? (int)(q / (cx * cy)) : 0;
            const double cell_x = (double)ix * c.dx - (((double)(r.nx - 1)) / 2.0 * c.dx);
            const double cell_y = (double)iy * c.dx - (((double)(r.ny - 1)) / 2.0 * c.dx);
            const double cell_z = d4 ? (double)iz * c.dx - (((double)(r.nz - 1)) / 2.0 * c.dx) : 0.0;
            double *o = rows.data() + (size_t)first[k] * 20;
            for (int64_t e = off[m]; e < off[m + 1]; ++e, o += 20) {
                const double *cen = el.data() + (size_t)e * per, *nor = cen + dim, *v = var.data() + (size_t)e * 10;
                o[0] = cen[0] + cell_tau;
                o[1] = cen[1] + cell_x;
                o[2] = cen[2] + cell_y;
                o[3] = d4 ? cen[3] + cell_z : 0.0;
                o[4] = t * nor[0];
                o[5] = t * nor[1];
                o[6] = t * nor[2];
                o[7] = d4 ? t * nor[3] : 0.0;
                o[8] = v[0];
                o[9] = v[1];
                o[10] = v[2];
                o[11] = v[3];
                o[12] = heos.temperature(v[3]);
                o[13] = heos.pressure(v[3]);
                for (int j = 0; j < 6; ++j) o[14 + j] = v[4 + j];
            }
        }
    });
}

double quantile(std::vector<double> v, double q) {  // nearest rank
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t k = (size_t)std::ceil(q * v.size());
    k = std::min(std::max<size_t>(k, 1), v.size());
    return v[k - 1];
}

}  // namespace

extern "C" vh_status vh_serve_run(const vh_serve_config *cfg, const double *eos_table, vh_serve_stats *stats,
                                  vh_serve_request *req) {
    if (!cfg || !stats || !req) return VH_ERR_ARG;
    const vh_serve_config &c = *cfg;
    if (c.n_requests < 1 || !(c.rate_per_s > 0) || c.coarse < 1 || c.max_resident < 1 || c.max_steps < c.coarse ||
        c.side2_min < 2 || c.side2_max < c.side2_min || c.side3_min < 2 || c.side3_max < c.side3_min ||
        c.nz3_min < 2 || c.nz3_max < c.nz3_min || !(c.dx > 0) || !(c.dt > 0) || !(c.t0 > 0) ||
        c.hbm_budget_bytes <= 0)
        return VH_ERR_ARG;
    if (c.collect_surface < 0 || c.collect_surface > 2) return VH_ERR_ARG;
    const int n = c.n_requests;
    const bool mixed = c.precision == VH_PREC_MIXED;
    // resident state, plus the freezeout slices: e's two, and the ten surface
    // variables' two when elements are built
    const int64_t bytes_per_cell = (mixed ? (3 * 15 + 3 * 4 + 3 * 4 + 4) * 4 : (3 * 15 + 3 * 4 + 4) * 8) +
                                   (c.collect_surface ? 2 * 8 : 0) + (c.collect_surface == 2 ? 20 * 8 : 0);

    // requests and their payload (generated before the clock starts)
    Rng rng{c.seed};
    double t_arr = 0;
    std::vector<Payload> pay(n);
    vh::HostEos heos;
    if (eos_table) heos.load(eos_table);
    size_t nmax = 1;
    for (int k = 0; k < n; ++k) {
        t_arr += -std::log(1.0 - rng.uniform()) / c.rate_per_s;
        vh_serve_request &r = req[k];
        r = vh_serve_request{};
        r.arrival_s = t_arr;
        const bool is3d = rng.uniform() < c.frac_3d;
        if (is3d) {
            r.nx = r.ny = rng.uniform_int(c.side3_min, c.side3_max);
            r.nz = rng.uniform_int(c.nz3_min, c.nz3_max);
        } else {
            r.nx = r.ny = rng.uniform_int(c.side2_min, c.side2_max);
            r.nz = 1;
        }
        r.glauber_seed = c.seed ^ (0x5DEECE66Dull * (uint64_t)(k + 1));
        const size_t N = (size_t)r.nx * r.ny * r.nz;
        nmax = std::max(nmax, N);
        vh_glauber_params gp{197, 6.38, 0.535, 4.2, -1.0, 10.0, 0.5, 4.0, 1.96, c.e0, c.e_floor, r.glauber_seed};
        pay[k].e.resize(N);
        vh_status s = vh_ic_mc_glauber(&gp, r.nx, r.ny, r.nz, c.dx, c.dx, c.dx, pay[k].e.data(), nullptr);
        if (s != VH_OK) return s;
        pay[k].p.resize(N);
        for (size_t m = 0; m < N; ++m) pay[k].p[m] = heos.pressure(pay[k].e[m]);
    }

    vh_batch_config bc{};
    bc.device = c.device;
    bc.capacity_cells = c.hbm_budget_bytes / bytes_per_cell;
    bc.max_events = c.max_resident;
    bc.max_steps = c.max_steps;
    bc.bulk = c.bulk;
    bc.eos_kind = eos_table ? VH_EOS_TABLE : VH_EOS_CONFORMAL;
    bc.eos_table = eos_table;
    bc.precision = c.precision;
    bc.fo_vars = c.collect_surface == 2;
    vh_batch *b = nullptr;
    vh_status st = vh_batch_create(&bc, &b);
    if (st != VH_OK) return st;
    const double eF = vh_batch_eos_energy_density(b, c.T_F_GeV / 0.197326938);

    std::vector<int> owner(c.max_resident, -1), strikes(c.max_resident, 0);
    std::vector<int64_t> counts(c.max_resident);
    std::vector<double> ebuf(nmax);
    // surface elements of one request's check: cube ids (per batch slot, kept
    // from the flags to the surface), counts, elements, variables, and the
    // rows handed to on_surface
    std::vector<std::vector<int64_t>> fo_ids(c.max_resident);
    std::vector<int64_t> fo_nflag(c.max_resident, 0);
    std::vector<int> fo_steps(c.max_resident, 0), leaving(c.max_resident, 0);
    std::vector<std::vector<int>> fo_ne(c.max_resident);
    std::vector<std::vector<double>> fo_el(c.max_resident), fo_var(c.max_resident);
    std::vector<int64_t> fo_tot(c.max_resident, 0);
    std::vector<double> fo_rows;
    std::vector<int64_t> fo_order, fo_bucket;
    // threads building the surface rows
    HostPool pool(c.collect_surface == 2 ? (int)std::min(7u, std::max(1u, std::thread::hardware_concurrency() / 2)) : 0);
    std::deque<int> queue;
    bool in_flight = false;  // the next check's steps are already enqueued
    int next = 0, finished = 0, resident = 0;
    int64_t resident_cells = 0;
    double occ_int = 0, occ_peak = 0, zone_updates = 0;
    const auto clock0 = std::chrono::steady_clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - clock0).count(); };
    double t_prev = 0;
    auto finish = [&](int id, int status) {
        const int k = owner[id];
        vh_serve_request &r = req[k];
        int active = 0, steps = 0;
        int64_t cells = 0;
        vh_batch_event_info(b, id, &active, &steps, &cells);
        r.steps = steps;
        r.status = status;
        if (status == VH_OK && vh_batch_download(b, id, 15, ebuf.data()) == VH_OK) {
            double sum = 0;
            for (int64_t m = 0; m < cells; ++m) sum += ebuf[m];
            r.e_sum = sum;
            if (c.digest) r.digest = state_digest(b, id, c.bulk ? 15 : 14, cells, ebuf);
        }
        vh_batch_remove(b, id);
        // the slot's surface arrays leave with the request (a large 3+1D
        // request's first check alone holds millions of elements)
        std::vector<int64_t>().swap(fo_ids[id]);
        std::vector<int>().swap(fo_ne[id]);
        std::vector<double>().swap(fo_el[id]);
        std::vector<double>().swap(fo_var[id]);
        r.done_s = now_s();
        owner[id] = -1;
        --resident;
        resident_cells -= cells;
        ++finished;
    };
    while (finished < n && st == VH_OK) {
        const double now = now_s();
        while (next < n && req[next].arrival_s <= now) queue.push_back(next++);
        // admission: first fit in arrival order; a request that cannot fit an
        // empty batch is refused outright
        for (auto it = queue.begin(); it != queue.end();) {
            const int k = *it;
            vh_serve_request &r = req[k];
            const int64_t N = (int64_t)r.nx * r.ny * r.nz;
            if (N > bc.capacity_cells) {
                r.status = VH_ERR_CAPACITY;
                r.admit_s = r.done_s = now;
                ++finished;
                it = queue.erase(it);
                continue;
            }
            vh_event_params ep{r.nx, r.ny, r.nz, c.dx, c.dx, c.dx, c.dt, c.t0, c.etabar, c.T_reg_GeV,
                               c.etas_Tk_GeV, c.etas_slope_lo, c.etas_slope_hi};
            int id = -1;
            // at rest without dissipative currents: only e and p are uploaded
            vh_status s = vh_batch_add(b, &ep, pay[k].e.data(), pay[k].p.data(), nullptr, nullptr, nullptr, nullptr,
                                       nullptr, nullptr, &id);
            if (s == VH_ERR_CAPACITY) {
                ++it;
                continue;
            }
            if (s != VH_OK) {
                st = s;
                break;
            }
            owner[id] = k;
            strikes[id] = 0;
            r.admit_s = now_s();
            ++resident;
            resident_cells += N;
            it = queue.erase(it);
        }
        if (st != VH_OK) break;
        if (resident == 0) {
            if (next < n) {
                const double wait = req[next].arrival_s - now_s();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
            }
            continue;
        }
        if (!in_flight) {
            st = vh_batch_steps(b, c.coarse);
            if (st != VH_OK) break;
            zone_updates += (double)resident_cells * c.coarse;
        }
        in_flight = false;
        // the check; a request whose state hit a NaN leaves as failed
        for (;;) {
            st = vh_batch_counts_above(b, eF, counts.data());
            if (st != VH_ERR_NAN) break;
            int site = 0, bad = -1;
            int64_t cell = 0;
            vh_batch_nan_info(b, &site, &bad, &cell);
            vh_batch_clear_error(b);
            if (bad < 0 || owner[bad] < 0) {
                st = VH_ERR_NAN;
                break;
            }
            finish(bad, VH_ERR_NAN);
        }
        if (st != VH_OK) break;
        // flags and the stop rule; a request admitted while the last steps
        // were in flight has not stepped yet and has its first check next time
        std::vector<int> chk;
        for (int id = 0; id < c.max_resident; ++id) {
            fo_nflag[id] = 0;
            leaving[id] = 0;
            if (owner[id] < 0) continue;
            int active = 0, steps = 0;
            vh_batch_event_info(b, id, &active, &steps, nullptr);
            if (steps == 0) continue;
            fo_steps[id] = steps;
            chk.push_back(id);
        }
        if (c.collect_surface && !chk.empty()) {
            // every checked request's cut cubes in one pass
            const size_t n = chk.size();
            std::vector<int64_t *> pid(n, nullptr);
            std::vector<int64_t> cap(n, 0), nflag(n, 0);
            for (size_t k = 0; k < n && c.collect_surface == 2; ++k) {
                const vh_serve_request &r = req[owner[chk[k]]];
                const int64_t ncube = (int64_t)(r.nx - 1) * (r.ny - 1) * (r.nz > 1 ? r.nz - 1 : 1);
                std::vector<int64_t> &ids = fo_ids[chk[k]];
                if ((int64_t)ids.size() < ncube) ids.resize(ncube);
                pid[k] = ids.data();
                cap[k] = ncube;
            }
            vh_status fs = vh_batch_fo_flags(b, (int)n, chk.data(), eF, pid.data(), cap.data(), nflag.data());
            if (fs != VH_OK && fs != VH_ERR_CAPACITY) {
                st = fs;
                break;
            }
            for (size_t k = 0; k < n; ++k) {
                req[owner[chk[k]]].fo_cubes += nflag[k];
                if (c.collect_surface == 2) fo_nflag[chk[k]] = nflag[k];
            }
        }
        for (int id : chk) {
            if (counts[id] == 0) ++strikes[id];
            leaving[id] = strikes[id] >= 3 || fo_steps[id] + c.coarse > c.max_steps;
        }
        if (st != VH_OK) break;
        // surfaces: the leaving requests' before their events leave the batch,
        // then the next check's steps go in flight while the others' are built
        // (vh_batch_fo_surfaces works from the flag's copies, on a stream of
        // its own)
        auto surfaces = [&](bool of_leaving) {
            std::vector<int> sel;
            for (int id = 0; id < c.max_resident; ++id)
                if (owner[id] >= 0 && fo_nflag[id] > 0 && (leaving[id] != 0) == of_leaving) sel.push_back(id);
            if (sel.empty()) return;
            vh_status fs = surface_pass(b, sel, c, fo_nflag, fo_ne, fo_el, fo_var, fo_tot);
            if (fs != VH_OK) {
                st = fs;
                return;
            }
            for (int id : sel) {
                vh_serve_request &r = req[owner[id]];
                surface_rows(r, c, heos, fo_steps[id], fo_ids[id], fo_nflag[id], fo_ne[id], fo_el[id], fo_var[id],
                             fo_tot[id], fo_rows, fo_order, fo_bucket, pool);
                const int64_t ne = (int64_t)(fo_rows.size() / 20);
                r.fo_elements += ne;
                if (c.on_surface && ne) c.on_surface(c.user, owner[id], ne, fo_rows.data());
            }
        };
        surfaces(true);
        if (st != VH_OK) break;
        for (int id = 0; id < c.max_resident; ++id)
            if (owner[id] >= 0 && leaving[id]) finish(id, VH_OK);
        if (resident > 0 && c.collect_surface == 2) {
            st = vh_batch_steps(b, c.coarse);
            if (st != VH_OK) break;
            zone_updates += (double)resident_cells * c.coarse;
            in_flight = true;
        }
        surfaces(false);
        if (st != VH_OK) break;
        const double t = now_s();
        const double occ = (double)(resident_cells * bytes_per_cell) / (double)c.hbm_budget_bytes;
        occ_int += occ * (t - t_prev);
        occ_peak = std::max(occ_peak, occ);
        t_prev = t;
    }
    const double wall = now_s();
    vh_batch_destroy(b);
    if (st != VH_OK) return st;
    std::vector<double> lat;
    int failed = 0;
    for (int k = 0; k < n; ++k) {
        if (req[k].status == VH_OK) lat.push_back(req[k].done_s - req[k].arrival_s);
        else ++failed;
    }
    stats->completed = (int)lat.size();
    stats->failed = failed;
    stats->wall_s = wall;
    stats->events_per_s = lat.size() / wall;
    stats->zone_upd_per_s = zone_updates / wall;
    stats->p50_s = quantile(lat, 0.50);
    stats->p99_s = quantile(lat, 0.99);
    double m = 0;
    for (double x : lat) m += x;
    stats->mean_s = lat.empty() ? 0.0 : m / lat.size();
    stats->occupancy_mean = wall > 0 ? occ_int / wall : 0.0;
    stats->occupancy_peak = occ_peak;
    return VH_OK;
}
