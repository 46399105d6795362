// Per-cell environment kernels, run at pre-loop and on environment events.
#pragma once
#include "cell_kernels.hpp"

namespace qhg {

constexpr int kBlock = 256;   // threads per block of the per-cell environment kernels

struct EnvPtrs {
    const double *alt, *lon, *lat, *water, *npp, *temp, *rain;
    const uint8_t *ice, *coastal;
};

struct CapParams {
    double water_factor, coastal_factor, coastal_min_lat, coastal_max_lat;
    double npp_min, npp_max, k_min, k_max, efficiency;
};

// NPPCapacity::recalculate, QHG4/actions/NPPCapacity.cpp:138-217, with the
// Miami NPP of QHG4/core/NPPCalcMiami.cpp:28-41 for the Oceania
// substitution (its exponentials through det_exp, so every cell's K is the
// same double on the device and on the host).
__global__ void __launch_bounds__(kBlock) k_capacity(Ctx x, EnvPtrs e, CapParams P) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < x.ncells;
         i += (int64_t)gridDim.x * kBlock) {
        double miami = 0;
        if (e.temp && e.rain) {
            const double nT = 3000.0 / (1 + det_exp(1.315 - 0.119 * e.temp[i]));
            const double nP = 3000.0 * (1 - det_exp(-0.000664 * e.rain[i]));
            miami = 0.000475 * ((nT < nP) ? nT : nP);
        }
        double cap;
        const double a = e.alt ? e.alt[i] : 0.0;
        if (a > 0) {
            const float A0 = 1500, A1 = 2500;
            double af = (a < A0) ? 1 : ((A1 - a) / (A1 - A0));
            if (af < 0) af = 0;
            const double npp = e.npp ? e.npp[i] : 0.0;
            double tn = npp;
            if (e.lon && e.lat && e.lon[i] > 115.0 && e.lat[i] > -12.0 && e.lon[i] < 150.0 && e.lat[i] < 1.0) {
                if (npp < P.npp_min) tn = miami;
            }
            tn *= af;
            if (tn < P.npp_min) cap = P.k_min;
            else if (tn > P.npp_max) cap = P.k_max;
            else cap = P.k_min + tn * P.k_max / (P.npp_max - P.npp_min);
            if (e.coastal && e.lat && e.coastal[i] && e.lat[i] > P.coastal_min_lat && e.lat[i] < P.coastal_max_lat)
                cap += P.coastal_factor * P.k_max;
            if (e.water) cap += e.water[i] * P.water_factor * P.k_max;
            if (cap > P.k_max) cap = P.k_max;
        } else {
            cap = 0;
        }
        x.K[i] = cap * P.efficiency;
    }
}

struct EvalParams {
    int32_t mode, cumulate, npts, first, last, pad;
    double weight;
    double px[QHG_MAX_PL_POINTS], pv[QHG_MAX_PL_POINTS];
};

// PolyLine::getVal, QHG4/utils/PolyLine.cpp:60-89
__device__ double polyline(const EvalParams &p, double X) {
    const int ns = p.npts - 1;
    if (X >= p.px[ns]) return p.pv[ns];
    int i = 0;
    while (i <= ns && X > p.px[i]) i++;
    if (i == 0) return p.pv[0];
    if (i <= ns) {
        const double a = (p.pv[i] - p.pv[i - 1]) / (p.px[i] - p.px[i - 1]);
        return p.pv[i - 1] + a * (X - p.px[i - 1]);
    }
    return p.pv[ns];
}

// SingleEvaluator::calcValues, QHG4/actions/SingleEvaluator.cpp:174-207
__global__ void __launch_bounds__(kBlock) k_eval_values(Ctx x, const double *in, double *tmp, EvalParams p) {
    for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < x.ncells;
         c += (int64_t)gridDim.x * kBlock) {
        double v = 0;
        if (x.ice == nullptr || !x.ice[c]) {
            v = (p.npts >= 2) ? polyline(p, (double)(float)in[c]) : in[c];
            v = (v > 0) ? v : 0;
        }
        tmp[c] = v;
    }
}

// SingleEvaluator::exchangeAndCumulate (SingleEvaluator.cpp:216-243) folded
// into MultiEvaluator's combination (MultiEvaluator.cpp:221-478); the last
// evaluator also takes the in-row prefix sum.
__global__ void __launch_bounds__(kBlock) k_eval_combine(Ctx x, const double *tmp, EvalParams p) {
    const int R = x.conn + 1;
    for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < x.ncells;
         c += (int64_t)gridDim.x * kBlock) {
        double *w = x.W + c * R;
        double s = tmp[c];
        double acc = s;
        for (int j = 0; j < R; j++) {
            double sj;
            if (j == 0) {
                sj = s;
            } else {
                const int32_t nb = x.nbr[c * x.conn + j - 1];
                double cw = (nb >= 0) ? tmp[nb] : 0.0;
                cw = (cw > 0) ? cw : 0;
                acc = p.cumulate ? acc + cw : cw;
                sj = acc;
            }
            double o;
            if (p.first) {
                o = (p.mode == QHG_ME_MUL_SIMPLE) ? 1.0
                  : (p.mode == QHG_ME_MAX_SIMPLE) ? -INFINITY
                  : (p.mode == QHG_ME_MIN_SIMPLE) ? INFINITY : 0.0;
            } else {
                o = w[j];
            }
            const double nv = sj * p.weight;
            switch (p.mode) {
            case QHG_ME_ADD_SIMPLE: o += nv; break;
            case QHG_ME_MUL_SIMPLE: o *= nv; break;
            case QHG_ME_MAX_SIMPLE: if (o < nv) o = nv; break;
            case QHG_ME_MIN_SIMPLE: if (o > nv) o = nv; break;
            }
            w[j] = o;
        }
        if (p.last && p.mode != QHG_ME_MAX_SIMPLE) {
            for (int j = 1; j < R; j++) w[j] += w[j - 1];
        }
    }
}

// The geography-event kill of OoAPop::updateEvent
// (QHG4/populations/OoAPop.cpp:164-196): every agent in a cell with
// altitude < 0 or ice is removed before the next step.  The buffer is cell
// sorted, so whole cells go.  Pass 1 (one block) lists the kept occupied
// cells of buffer b, in order, in occ_dst with dense segment starts, their
// old starts in src_s, and the resulting counts in the other parity's
// status; pass 2 moves the kept agents into the other buffer, which then
// takes the current buffer's role.
constexpr int kKillBlock = 1024;

__global__ void __launch_bounds__(kKillBlock) k_env_kill(const OccCell *occ, int64_t nocc, const double *alt,
                                                         const uint8_t *ice, OccCell *occ_dst, uint32_t *src_s,
                                                         DevStatus *st, int b) {
    constexpr int kWaves = kKillBlock / kWave;
    __shared__ uint32_t wkeep[kWaves], wslots[kWaves];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    int64_t out = 0;      // kept cells so far
    uint32_t slots = 0;   // their agents
    for (int64_t base = 0; base < nocc; base += kKillBlock) {
        const int64_t i = base + threadIdx.x;
        OccCell e{0, 0, 0, 0};
        bool keep = false;
        if (i < nocc) {
            e = occ[i];
            keep = !((alt && alt[e.c] < 0) || (ice && ice[e.c] > 0));
        }
        const uint32_t v = keep ? e.n : 0u;
        uint32_t incl = v;
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d, kWave);
            if (lane >= d) incl += y;
        }
        const uint64_t m = __ballot(keep);
        const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wkeep[w] = (uint32_t)__popcll(m);
        if (lane == kWave - 1) wslots[w] = incl;
        __syncthreads();
        uint32_t pre = 0, total = 0, spre = 0, stotal = 0;
        for (int k = 0; k < kWaves; k++) {
            if (k < w) {
                pre += wkeep[k];
                spre += wslots[k];
            }
            total += wkeep[k];
            stotal += wslots[k];
        }
        if (keep) {
            const int64_t o = out + pre + below;
            occ_dst[o] = OccCell{e.c, slots + spre + incl - v, e.n, 0u};
            src_s[o] = e.s;
        }
        out += total;
        slots += stotal;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int nb = 1 - b;
        st->n_occ[nb] = out;
        st->n_alive[nb] = slots;
        st->max_id[nb] = st->max_id[b];
        st->gpool[nb] = st->gpool[b];
    }
}

// Pass 2: one wave per kept cell copies its agents.  The last pairing no
// longer holds, so mate ids read back as unpaired (-3) until the next step
// pairs again.
__global__ void __launch_bounds__(kBlock) k_env_move(AgentBuf from, AgentBuf to, const OccCell *occ,
                                                     const uint32_t *src_s, const DevStatus *st, int nb) {
    const int64_t nocc = st->n_occ[nb];
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
    for (int64_t k = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave; k < nocc; k += nwaves) {
        const OccCell e = occ[k];
        const uint32_t s = src_s[k];
        for (uint32_t j = lane; j < e.n; j += kWave) {
            const uint32_t f = s + j, t = e.s + j;
            to.id[t] = from.id[f];
            to.mate_id[t] = -3;
            to.life[t] = from.life[f];
            to.gender[t] = from.gender[f];
            to.birth[t] = from.birth[f];
            to.age[t] = from.age[f];
            to.lastb[t] = from.lastb[f];
            to.nbabies[t] = from.nbabies[f];
            if (to.ghandle) to.ghandle[t] = from.ghandle[f];
        }
    }
}

}  // namespace qhg

namespace qhg {

// The output function of SplitMix64 (Steele, Lea, Flood: "Fast splittable
// pseudorandom number generators", OOPSLA 2014): a bijective 64-bit mix.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Digest of a population's state: every live agent's fields (id, cell,
// life state, gender, birth time, age, last birth, stored partner id,
// number of babies, genome words) are folded into one 64-bit hash, and the
// hashes are summed in two independent ways, so the digest does not depend
// on the order the agents are stored in.  One wave per occupied cell in
// grid-stride order; out[0..1] (zeroed by the caller) receive the sums.
__global__ void __launch_bounds__(kBlock) k_state_digest(AgentBuf a, const OccCell *occ, int64_t nocc,
                                                         const uint64_t *gpool, int32_t gwords,
                                                         unsigned long long *out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nwaves = ((int64_t)gridDim.x * kBlock) / kWave;
    uint64_t s0 = 0, s1 = 0;
    for (int64_t e = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave; e < nocc; e += nwaves) {
        const OccCell oc = occ[e];
        for (uint32_t i = oc.s + lane; i < oc.s + oc.n; i += kWave) {
            uint64_t h = 0x9E3779B97F4A7C15ull;
            h = mix64(h ^ (uint64_t)a.id[i]);
            h = mix64(h ^ (uint64_t)(uint32_t)oc.c);
            h = mix64(h ^ ((uint64_t)a.life[i] | ((uint64_t)a.gender[i] << 8)));
            h = mix64(h ^ (uint64_t)__double_as_longlong(a.birth[i]));
            h = mix64(h ^ (uint64_t)__double_as_longlong(a.age[i]));
            h = mix64(h ^ (uint64_t)__double_as_longlong(a.lastb[i]));
            h = mix64(h ^ (uint64_t)a.mate_id[i]);
            h = mix64(h ^ (uint64_t)(uint32_t)a.nbabies[i]);
            if (gwords > 0) {
                const uint64_t *g = gpool + (size_t)a.ghandle[i] * (size_t)gwords;
                for (int32_t k = 0; k < gwords; k++) h = mix64(h ^ g[k]);
            }
            s0 += h;
            s1 += mix64(h ^ 0xD6E8FEB86659FD93ull);
        }
    }
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) {
        s0 += (uint64_t)__shfl_xor((unsigned long long)s0, d, kWave);
        s1 += (uint64_t)__shfl_xor((unsigned long long)s1, d, kWave);
    }
    if (lane == 0 && (s0 | s1)) {
        atomicAdd(&out[0], (unsigned long long)s0);
        atomicAdd(&out[1], (unsigned long long)s1);
    }
}

}  // namespace qhg
