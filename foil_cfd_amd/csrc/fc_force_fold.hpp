// Force EMA and trailing-average statistics of the solver's momentum-exchange
// samples (reference LBMSolver: fold lbm_solver.cpp:633-654, readout
// lbm_solver.cpp:1414-1509, ring size lbm_solver.cpp:120).
//
// fold(): the EMA weight covers the steps since the previous folded sample,
// alpha = 1 - exp(-dSteps / windowSteps), so the effective window is the
// configured number of flow-throughs whatever the batch length; the first
// sample after a (re)start seeds the EMA. Every raw sample also enters a
// 256-entry ring.
//
// readout(): coefficients of the EMA; when the caller's gate is open, the
// mean / min / max / median of Cl, Cd and L/D over the ring samples taken at
// or after one flow-through (the start-up transient is dropped). Sums run in
// FP32 from the newest sample to the oldest, as the reference's do.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/foilcfd.h"
#include "fc_units.hpp"

namespace fc {

class ForceFold {
public:
    static constexpr int kRingSize = 256;

    // Cold start: EMA unseeded and the ring emptied.
    void clear() {
        emaFx_ = emaFy_ = 0.0f;
        seeded_ = false;
        head_ = count_ = 0;
    }
    // Warm restart: the EMA restarts from the next sample; the ring keeps
    // its samples (the reference clears it on cold starts only).
    void unseed() { seeded_ = false; }
    bool seeded() const { return seeded_; }
    float emaFx() const { return emaFx_; }
    float emaFy() const { return emaFy_; }

    void fold(long long atStep, float dSteps, float windowSteps, float fx, float fy) {
        const float alpha = 1.0f - std::exp(-dSteps / std::max(1.0f, windowSteps));
        if (!seeded_) {
            emaFx_ = fx;
            emaFy_ = fy;
            seeded_ = true;
        } else {
            emaFx_ += alpha * (fx - emaFx_);
            emaFy_ += alpha * (fy - emaFy_);
        }
        ring_[head_] = Sample{atStep, fx, fy};
        head_ = (head_ + 1) % kRingSize;
        if (count_ < kRingSize) ++count_;
    }

    // r->flow_throughs is left to the caller.
    void readout(const Scaling& s, int span, long long oneFlowThroughSteps, bool gateOpen,
                 fc_force_readout* r) const {
        r->cd = s.forceToCoefficient(emaFx_, span);
        r->cl = s.forceToCoefficient(emaFy_, span);
        r->lift_to_drag = (std::fabs(r->cd) > 1e-6f) ? r->cl / r->cd : 0.0f;
        r->valid = (seeded_ && gateOpen) ? 1 : 0;
        r->cl_avg = r->cl_min = r->cl_max = r->cl_median = 0.0f;
        r->cd_avg = r->cd_min = r->cd_max = r->cd_median = 0.0f;
        r->ld_avg = r->ld_min = r->ld_max = r->ld_median = 0.0f;
        if (!r->valid || count_ == 0) return;
        std::vector<float> cl, cd, ld;
        cl.reserve(count_);
        cd.reserve(count_);
        ld.reserve(count_);
        for (int i = 0; i < count_; ++i) {  // newest first
            const Sample& smp = ring_[(head_ - 1 - i + kRingSize) % kRingSize];
            if (smp.step < oneFlowThroughSteps) continue;
            const float c1 = s.forceToCoefficient(smp.fy, span);
            const float c2 = s.forceToCoefficient(smp.fx, span);
            cl.push_back(c1);
            cd.push_back(c2);
            ld.push_back((std::fabs(c2) > 1e-6f) ? c1 / c2 : 0.0f);
        }
        if (cl.empty()) return;
        stats(cl, r->cl_avg, r->cl_min, r->cl_max, r->cl_median);
        stats(cd, r->cd_avg, r->cd_min, r->cd_max, r->cd_median);
        stats(ld, r->ld_avg, r->ld_min, r->ld_max, r->ld_median);
    }

private:
    struct Sample {
        long long step = 0;
        float fx = 0, fy = 0;
    };

    static void stats(std::vector<float> v, float& avg, float& mn, float& mx, float& med) {
        float sum = 0.0f;
        for (float x : v) sum += x;
        avg = sum / static_cast<float>(v.size());
        mn = *std::min_element(v.begin(), v.end());
        mx = *std::max_element(v.begin(), v.end());
        std::sort(v.begin(), v.end());
        const std::size_t m = v.size() / 2;
        med = (v.size() % 2 == 0) ? 0.5f * (v[m - 1] + v[m]) : v[m];
    }

    float emaFx_ = 0.0f, emaFy_ = 0.0f;
    bool seeded_ = false;
    Sample ring_[kRingSize]{};
    int head_ = 0, count_ = 0;
};

}  // namespace fc
