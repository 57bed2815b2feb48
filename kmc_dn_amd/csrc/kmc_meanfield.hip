// Mean-field estimators of the Go library, one wave per network.
//
// probSimulate (goSimulation/probabilitySimulation.go:53-157, behind
// wrapperSimulateProbability, simulationWrapper.go:235-250) integrates
// fractional occupations deterministically.  Each step:
//   * site energies from the occupations (:82-90), FP64;
//   * every pair's rate t_ij = p_ij * nu * min(1, exp(-dE/kT)) * tc_ij with the
//     occupation factor p_ij of probTransitionPossible (:40-50) (calcProbTransitions
//     :8-38), and their total;
//   * a step that moves each occupation by its net flow over the total, capped so
//     that none leaves [0, 1] (:96-110), the clock advancing 0.98 x that step;
//   * occupation, electrode occupation and net traffic moved by rate x cap / total.
// Lanes own sites s = lane + 64c.  Each lane sums its site's row (rates out) and
// column (rates in), so a step needs only the wave's total and the wave's
// minimum cap: no S x S table and no cross-lane work per pair.  Sums run in a
// different order than the reference's row-major loops, so results agree to
// rounding, not bit for bit.
//
// calcAverageCurrent (goSimulation/simulation.go:671-736, the second half of
// wrapperSimulateCombined) evaluates the same rates once at a time-averaged
// occupation, with the reference's FP32 site energies and energy differences
// (the wrapper casts its inputs to float32), and returns each electrode's net
// inflow over the total rate divided by the reference's time step 0.98 / total.
#include "kmc_device.h"

namespace kmc {

namespace {

// probTransitionPossible (probabilitySimulation.go:40-50)
__device__ __forceinline__ double occupation_factor(int i, int j, int N, double oi, double oj) {
  if (i >= N && j >= N) return 0.0;
  if (i >= N) return 1.0 - oj;
  if (j >= N) return oi;
  return (1.0 - oj) * oi;
}

__device__ __forceinline__ double wave_sum_f64(double x) {
  const double incl = wave_scan_f64(x);
  return readlane_f64(incl, 63);
}

__device__ __forceinline__ double wave_min_f64(double x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double y = __shfl_xor(x, o, 64);
    x = y < x ? y : x;
  }
  return x;
}

}  // namespace

template <int C>
__global__ void __launch_bounds__(64) kmc_meanfield_kernel(MeanfieldParams p) {
  __shared__ double occ_l[64 * C], se_l[64 * C];
  const int k = blockIdx.x;
  if (k >= p.n) return;
  const int lane = threadIdx.x;
  const LayoutDev L = p.layouts[p.layout_of[k]];
  const EsetDev es = p.esets[p.eset_of[k]];
  const int N = L.N, S = L.S;
  const double* d = L.dist64;
  const double* tc = L.tc64;
  const double nu = L.nu, kT = L.kT, I0R = L.I0R;

  double occ[C], eo[C], avg[C], se[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    occ[c] = s < N ? p.occ[(size_t)k * NMAX + s] : 0.0;
    occ_l[s] = occ[c];
    eo[c] = 0.0;
    avg[c] = 0.0;
    se[c] = 0.0;
  }
  double time = 0.0;
  for (int64_t step = 0; step < p.steps; ++step) {
    wave_lds_sync();
    // site energies (:82-90): Coulomb sum over the other acceptors, ascending j
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int s = lane + 64 * c;
      if (s < N) {
        double ai = 0.0;
        for (int j = 0; j < N; ++j)
          if (j != s) ai += (1.0 - occ_l[j]) / d[(size_t)s * S + j];
        se[c] = es.E[s] - I0R * ai;
      } else if (s < S) {
        se[c] = es.V[s - N];
      }
      se_l[s] = se[c];
    }
    wave_lds_sync();
    // rates out of (row) and into (column) each lane's site
    double row[C], col[C];
#pragma unroll
    for (int c = 0; c < C; ++c) row[c] = col[c] = 0.0;
    for (int j = 0; j < S; ++j) {
      const double oj = j < N ? occ_l[j] : 0.0, sej = se_l[j];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int s = lane + 64 * c;
        if (s >= S) continue;
        const bool acc = s < N && j < N;
        const double cd = acc ? I0R / d[(size_t)s * S + j] : 0.0;
        const double dout = sej - se[c] - cd;   // s -> j
        const double din = se[c] - sej - cd;    // j -> s
        const double mout = dout > 0 ? nu * exp(-dout / kT) : nu;
        const double min_ = din > 0 ? nu * exp(-din / kT) : nu;
        row[c] += occupation_factor(s, j, N, occ[c], oj) * mout * tc[(size_t)s * S + j];
        col[c] += occupation_factor(j, s, N, oj, occ[c]) * min_ * tc[(size_t)j * S + s];
      }
    }
    double part = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c)
      if (lane + 64 * c < S) part += row[c];
    const double tot = wave_sum_f64(part);
    // the largest step that keeps every occupation in [0, 1] (:96-110)
    double cap = 1.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int s = lane + 64 * c;
      if (s >= N) continue;
      const double dr = (col[c] - row[c]) / tot;
      const double nv = occ[c] + dr;
      if (nv < 0) {
        const double r = occ[c] / -dr;
        cap = r < cap ? r : cap;
      }
      if (nv > 1) {
        const double r = (1.0 - occ[c]) / dr;
        cap = r < cap ? r : cap;
      }
    }
    const double max_change = wave_min_f64(cap);
    const double time_step = 0.98 * max_change / tot;
    time += time_step;
    if (p.avg) {
#pragma unroll
      for (int c = 0; c < C; ++c) avg[c] += occ[c] * time_step;
    }
    const double f = max_change / tot;
    if (p.traffic) {   // net traffic of every pair of the lane's row
      double* tr = p.traffic + (size_t)k * p.tstride;
      for (int j = 0; j < S; ++j) {
        const double oj = j < N ? occ_l[j] : 0.0, sej = se_l[j];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const int s = lane + 64 * c;
          if (s >= S || (s >= N && j >= N)) continue;
          const bool acc = s < N && j < N;
          const double cd = acc ? I0R / d[(size_t)s * S + j] : 0.0;
          const double dout = sej - se[c] - cd, din = se[c] - sej - cd;
          const double mout = dout > 0 ? nu * exp(-dout / kT) : nu;
          const double min_ = din > 0 ? nu * exp(-din / kT) : nu;
          const double out = occupation_factor(s, j, N, occ[c], oj) * mout * tc[(size_t)s * S + j];
          const double in = occupation_factor(j, s, N, oj, occ[c]) * min_ * tc[(size_t)j * S + s];
          tr[(size_t)s * S + j] += (out - in) * f;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int s = lane + 64 * c;
      if (s < N) {
        double o = occ[c] + (col[c] - row[c]) * f;
        o = o < 0 ? 0.0 : (o > 1 ? 1.0 : o);   // (:146-155)
        occ[c] = o;
      } else if (s < S) {
        eo[c] += (col[c] - row[c]) * f;
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int c = 0; c < C; ++c) occ_l[lane + 64 * c] = occ[c];
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    if (s < N) {
      p.occ[(size_t)k * NMAX + s] = occ[c];
      if (p.se) p.se[(size_t)k * NMAX + s] = se[c];
      if (p.avg) p.avg[(size_t)k * NMAX + s] = avg[c];
    } else if (s < S) {
      p.eo[(size_t)k * PMAX + (s - N)] = eo[c];
    }
  }
  if (lane == 0 && p.time) p.time[k] = time;
}

// calcAverageCurrent (simulation.go:671-736) at the occupations p.occ; writes the
// currents into p.eo.  Site energies and energy differences are FP32 as in the
// reference (its inputs are float32 copies); rates and sums are FP64.
template <int C>
__global__ void __launch_bounds__(64) kmc_avg_current_kernel(MeanfieldParams p) {
  __shared__ double occ_l[64 * C];
  __shared__ float se_l[64 * C];
  const int k = blockIdx.x;
  if (k >= p.n) return;
  const int lane = threadIdx.x;
  const LayoutDev L = p.layouts[p.layout_of[k]];
  const EsetDev es = p.esets[p.eset_of[k]];
  const int N = L.N, S = L.S;
  const double* d = L.dist64;
  const double* tc = L.tc64;
  const float nu = (float)L.nu, kT = (float)L.kT, I0 = (float)L.I0, R = (float)L.R;
  double occ[C];
  float se[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    occ[c] = s < N ? p.occ[(size_t)k * NMAX + s] : 0.0;
    occ_l[s] = occ[c];
  }
  wave_lds_sync();
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    se[c] = 0.f;
    if (s < N) {
      double ai = 0.0;
      for (int j = 0; j < N; ++j)
        if (j != s) ai += (1.0 - occ_l[j]) / (double)(float)d[(size_t)s * S + j];
      se[c] = (float)es.E[s] - I0 * R * (float)ai;
    } else if (s < S) {
      se[c] = (float)es.V[s - N];
    }
    se_l[s] = se[c];
  }
  wave_lds_sync();
  double row[C], col[C];
#pragma unroll
  for (int c = 0; c < C; ++c) row[c] = col[c] = 0.0;
  for (int j = 0; j < S; ++j) {
    const double oj = j < N ? occ_l[j] : 0.0;
    const float sej = se_l[j];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int s = lane + 64 * c;
      if (s >= S) continue;
      const bool acc = s < N && j < N;
      const float cd = acc ? I0 * R / (float)d[(size_t)s * S + j] : 0.f;
      const float dout = acc ? sej - se[c] - cd : sej - se[c];
      const float din = acc ? se[c] - sej - cd : se[c] - sej;
      const double mout = dout > 0 ? (double)nu * exp((double)(-dout / kT)) : (double)nu;
      const double min_ = din > 0 ? (double)nu * exp((double)(-din / kT)) : (double)nu;
      row[c] += occupation_factor(s, j, N, occ[c], oj) * mout * (double)(float)tc[(size_t)s * S + j];
      col[c] += occupation_factor(j, s, N, oj, occ[c]) * min_ * (double)(float)tc[(size_t)j * S + s];
    }
  }
  double part = 0.0;
#pragma unroll
  for (int c = 0; c < C; ++c)
    if (lane + 64 * c < S) part += row[c];
  const double tot = wave_sum_f64(part);
  const double time_step = 0.98 / tot;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    if (s >= N && s < S) p.eo[(size_t)k * PMAX + (s - N)] = ((col[c] - row[c]) / tot) / time_step;
  }
}

template __global__ void kmc_meanfield_kernel<1>(MeanfieldParams);
template __global__ void kmc_meanfield_kernel<2>(MeanfieldParams);
template __global__ void kmc_meanfield_kernel<3>(MeanfieldParams);
template __global__ void kmc_avg_current_kernel<1>(MeanfieldParams);
template __global__ void kmc_avg_current_kernel<2>(MeanfieldParams);
template __global__ void kmc_avg_current_kernel<3>(MeanfieldParams);

}  // namespace kmc
