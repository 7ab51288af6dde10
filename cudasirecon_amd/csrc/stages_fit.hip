// The k0 search's peak refinement and the k0/modamp line searches on the
// device (findk0, gpuFunctionsImpl.cu:316-370; fitk0andmodamps, 874-1020;
// findrealspacemodamp's result, 1128-1147).  One wavefront per SIM direction
// runs the same walk the host would: each walker step replays a direction's
// walk from its start over the candidates evaluated so far and either
// finishes it (the parabola through its last three points) or asks for the
// next kFitLook steps from the first value it misses.  The float and double
// expressions are the host's, in the same order, and the angles and candidate
// vectors come from the same bit-reproducible cos/sin/atan2 (dtrig.hpp), so
// the device walks evaluate the host walks' candidates.
#include "dtrig.hpp"
#include "stages.hpp"

namespace simr {

namespace {

__device__ float fit_parabola(float a1, float a2, float a3) {
#pragma clang fp contract(off)
  const float slope = (float)(0.5 * (double)(a3 - a1));
  const float curve = (a3 + a1) - 2 * a2;
  if (curve == 0) return 0.0f;
  const float peak = -slope / curve;
  if (peak > 1.5f || peak < -1.5f) return 0.0f;
  return peak;
}

template <typename U>
__device__ U fit_xyparabola(U x1, U y1, U x2, U y2, U x3, U y3) {
#pragma clang fp contract(off)
  if (x1 == x2 || x2 == x3 || x3 == x1) return 0;
  const U xbar1 = (U)(0.5 * (double)(x1 + x2));
  const U xbar2 = (U)(0.5 * (double)(x2 + x3));
  const U slope1 = (y2 - y1) / (x2 - x1);
  const U slope2 = (y3 - y2) / (x3 - x2);
  const U curve = (slope2 - slope1) / (xbar2 - xbar1);
  if (curve == 0) return 0;
  return xbar2 - slope2 / curve;
}

// |modamp|^2 of one evaluation (xr, xi, sx, sy), as the host forms it
__device__ double fit_amp2(const double* e, int fp32) {
#pragma clang fp contract(off)
  if (fp32) {
    const float xr = (float)e[0], xi = (float)e[1], sx = (float)e[2];
    const float mx = xr / sx, my = xi / sx;
    const float a2 = mx * mx + my * my;
    return (double)a2;
  }
  const double mx = e[0] / e[2], my = e[1] / e[2];
  return mx * mx + my * my;
}

// One direction's walker state while a walker step runs: one wavefront per
// direction, every lane holding the same scalar state (the logic is
// replicated across the lanes; only the cache lookups are spread over them),
// the evaluated keys and this round's requests in LDS.
struct Walker {
  FitDir* f;
  float* ka;  // evaluated (angle, mag) keys [kFitCache]
  float* km;
  double* amp2;  // their |modamp|^2
  int ncache;
  float* ra;  // this round's requests [kFitReq]
  float* rm;
  int nreq;
  int lane;
};

__device__ void walker_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// index of the evaluated key (a, m), or -1: the lanes test 64 keys at a time
__device__ int walker_find(const Walker& w, float a, float m) {
  const unsigned ua = __float_as_uint(a), um = __float_as_uint(m);
  for (int base = 0; base < w.ncache; base += 64) {
    const int i = base + w.lane;
    const int ic = i < w.ncache ? i : 0;  // an index inside the array for every lane
    const bool hit = i < w.ncache && __float_as_uint(w.ka[ic]) == ua && __float_as_uint(w.km[ic]) == um;
    const unsigned long long b = __ballot(hit);
    if (b != 0ull) return base + __ffsll((long long)b) - 1;
  }
  return -1;
}

// Adds (a, m) to the round's requests unless it is evaluated or already
// requested.
__device__ void walker_request(Walker& w, float a, float m) {
  if (walker_find(w, a, m) >= 0) return;
  const unsigned ua = __float_as_uint(a), um = __float_as_uint(m);
  // the requests so far (at most kFitReq < 64), one per lane
  const int li = w.lane < kFitReq ? w.lane : 0;  // an index inside the array for every lane
  const bool dup = w.lane < w.nreq && __float_as_uint(w.ra[li]) == ua && __float_as_uint(w.rm[li]) == um;
  if (__ballot(dup) != 0ull) return;
  if (w.nreq < kFitReq) {
    if (w.lane == 0) {
      w.ra[w.nreq] = a;
      w.rm[w.nreq] = m;
    }
    walker_sync();
    ++w.nreq;
  }
}

// kFitLook steps of a walk from `a` (angle walks vary the angle at a fixed
// magnitude, magnitude walks the magnitude), by the float accumulation of
// the step.
__device__ void walker_request_walk(Walker& w, float a, float other, bool is_angle, float delta, int sign) {
#pragma clang fp contract(off)
  for (int m = 0; m < kFitLook; ++m) {
    if (is_angle)
      walker_request(w, a, other);
    else
      walker_request(w, other, a);
    a = sign > 0 ? a + delta : a - delta;
  }
}

// One walk of fitk0andmodamps (gpuFunctionsImpl.cu:912-983) replayed over the
// cache: from `start` step by `delta` while the modamp magnitude grows, then
// the parabola through the last three points.  true: *out and *nevals set;
// false: the next kFitLook steps from the first missing value are requested.
__device__ bool walker_replay(Walker& w, bool is_angle, float start, float other, float delta, int fp32, float* out,
                              int* nevals) {
#pragma clang fp contract(off)
  int n = 0;
  int sign = +1;
  auto val = [&](float a, double* amp) {
    const int i = is_angle ? walker_find(w, a, other) : walker_find(w, other, a);
    if (i < 0) {
      walker_request_walk(w, a, other, is_angle, delta, sign);
      return false;
    }
    *amp = w.amp2[i];
    ++n;
    return true;
  };
  float x2 = start;
  double amp2, amp3, amp1 = 0;
  if (!val(start, &amp2)) return false;
  float a = start + delta;
  float x3 = a;
  if (!val(a, &amp3)) return false;
  float x1 = 0;
  if (amp3 > amp2) {
    while (amp3 > amp2) {
      amp1 = amp2;
      x1 = x2;
      amp2 = amp3;
      x2 = x3;
      a += delta;
      x3 = a;
      if (!val(a, &amp3)) return false;
    }
  } else {
    a = start;
    const double t = amp3;
    amp3 = amp2;
    amp2 = t;
    const float tx = x3;
    x3 = x2;
    x2 = tx;
    sign = -1;
    while (amp3 > amp2) {
      amp1 = amp2;
      x1 = x2;
      amp2 = amp3;
      x2 = x3;
      a -= delta;
      x3 = a;
      if (!val(a, &amp3)) return false;
    }
  }
  if (!fp32)
    *out = (float)fit_xyparabola<double>(x1, amp1, x2, amp2, x3, amp3);
  else
    *out = fit_xyparabola<float>(x1, (float)amp1, x2, (float)amp2, x3, (float)amp3);
  *nevals = n;
  return true;
}

// The order-fo2 modamp and correlation of the final evaluation
// (findrealspacemodamp's result, 1128-1147).
__device__ void fit_modamp(const double* e, int fp32, float2* m, float* cc) {
#pragma clang fp contract(off)
  if (fp32) {
    const float xr = (float)e[0], xi = (float)e[1], sx = (float)e[2], sy = (float)e[3];
    const float mx = xr / sx, my = xi / sx;
    *m = make_float2(mx, my);
    *cc = sqrtf((xr * xr + xi * xi) / (sx * sy));
  } else {
    const double mx = e[0] / e[2], my = e[1] / e[2];
    *m = make_float2((float)mx, (float)my);
    *cc = (float)sqrt((e[0] * e[0] + e[1] * e[1]) / (e[2] * e[3]));
  }
}

__global__ void k_fit_start(FitState* fs, FitArgs a) {
#pragma clang fp contract(off)
  const int d = threadIdx.x;
  if (d == 0) {
    fs->ncand = 0;
    fs->ndone = 0;
    fs->rounds = 0;
    fs->overflow = 0;
  }
  if (d >= a.ndirs) return;
  FitDir& f = fs->d[d];
  const float* v = fs->peak_vals + d * 8;
  const int idx = fs->peak_idx[d];
  const int xc = idx % a.nx, yc = idx / a.nx;
  float2 pk;
  pk.x = fit_parabola(v[1], v[0], v[2]) + (float)xc;
  pk.y = fit_parabola(v[3], v[0], v[4]) + (float)yc;
  const float2 old = a.guess[d];
  if (old.x / a.dkx < pk.x - (float)(a.nx / 2)) pk.x -= (float)a.nx;
  if (old.x / a.dkx > pk.x + (float)(a.nx / 2)) pk.x += (float)a.nx;
  if (old.y / a.dky < pk.y - (float)(a.ny / 2)) pk.y -= (float)a.ny;
  if (old.y / a.dky > pk.y + (float)(a.ny / 2)) pk.y += (float)a.ny;
  pk.x *= a.dkx / (float)a.fo2;
  pk.y *= a.dky / (float)a.fo2;
  f.k0 = pk;
  f.k0mag = sqrtf(pk.x * pk.x + pk.y * pk.y);
  f.k0angle = dtrig::fatan2(pk.y, pk.x);
  f.kov = make_float2(f.k0mag * dtrig::fcos(f.k0angle), f.k0mag * dtrig::fsin(f.k0angle));
  f.angle = f.mag = 0.f;
  f.phase = kFitAngle;
  f.started = 0;
  f.nreq = 0;
  f.cand0 = 0;
  f.n_angle_evals = f.n_mag_evals = f.n_reductions = 0;
  f.ncache = 0;
}

// One wavefront per direction (blockDim = 64 * ndirs).
__global__ __launch_bounds__(64 * kMaxDirs) void k_fit_walk(FitState* fs, FitArgs a, float2* amps, float* corr) {
#pragma clang fp contract(off)
  __shared__ float ska[kMaxDirs][kFitCache], skm[kMaxDirs][kFitCache];
  __shared__ double samp[kMaxDirs][kFitCache];
  __shared__ float sra[kMaxDirs][kFitReq], srm[kMaxDirs][kFitReq];
  __shared__ int snreq[kMaxDirs], sdone[kMaxDirs], sover[kMaxDirs];
  const int d = threadIdx.x / 64, lane = threadIdx.x % 64;
  const bool active = d < a.ndirs;  // uniform over each wavefront
  FitDir& f = fs->d[active ? d : 0];
  int ncache = 0;
  if (active) {
    // the keys evaluated so far, then the last round's sums of this
    // direction's requests
    ncache = f.ncache;
    const int nprev = f.nreq, cand0 = f.cand0;
    const int nadd = nprev < kFitCache - ncache ? nprev : kFitCache - ncache;
    for (int i = lane; i < ncache; i += 64) {
      ska[d][i] = f.ca[i];
      skm[d][i] = f.cm[i];
      samp[d][i] = fit_amp2(f.cev[i], a.fp32);
    }
    for (int i = lane; i < nadd; i += 64) {
      const int k = ncache + i;
      const float ra = f.req_a[i], rm = f.req_m[i];
      ska[d][k] = ra;
      skm[d][k] = rm;
      f.ca[k] = ra;
      f.cm[k] = rm;
      double e[4];
      for (int q = 0; q < 4; ++q) e[q] = fs->ev[cand0 + i][q];
      for (int q = 0; q < 4; ++q) f.cev[k][q] = e[q];
      samp[d][k] = fit_amp2(e, a.fp32);
    }
    if (lane == 0) sover[d] = nadd < nprev ? 1 : 0;
    ncache += nadd;
  }
  __syncthreads();  // the merged sums are read back below by other lanes
  if (active) {
    Walker w{&f, ska[d], skm[d], samp[d], ncache, sra[d], srm[d], 0, lane};
    int phase = f.phase, started = f.started;
    float angle = f.angle, mag = f.mag;
    const float k0angle = f.k0angle, k0mag = f.k0mag;
    int nang = f.n_angle_evals, nmag = f.n_mag_evals, nred = f.n_reductions;
    float2 kov = f.kov;
    for (int guard = 0; guard < 4; ++guard) {
      if (phase == kFitAngle || phase == kFitMag) {
        const bool ang = phase == kFitAngle;
        const float s0 = ang ? k0angle : k0mag;
        const float other = ang ? k0mag : angle;
        const float delta = ang ? a.deltaangle : a.deltamag;
        if (!started) {
          // the start and the first kFitLook steps both ways in one batch
          started = 1;
          if (ang)
            walker_request(w, s0, other);
          else
            walker_request(w, other, s0);
          walker_request_walk(w, s0 + delta, other, ang, delta, +1);
          walker_request_walk(w, s0 - delta, other, ang, delta, -1);
          if (w.nreq > 0) break;
        }
        float r;
        int ne = 0;
        if (!walker_replay(w, ang, s0, other, delta, a.fp32, &r, &ne)) break;
        if (ang) {
          angle = r;
          nang += ne;
        } else {
          mag = r;
          nmag += ne;
        }
        nred += ne;
        phase += 1;
        started = 0;
      } else if (phase == kFitFinal) {
        const int i = walker_find(w, angle, mag);
        if (i < 0) {
          walker_request(w, angle, mag);
          break;
        }
        // getmodamp at the fitted angle and magnitude: the fitted k0 and its modamp
        kov = make_float2(mag * dtrig::fcos(angle), mag * dtrig::fsin(angle));
        float2 m;
        float cc;
        fit_modamp(f.cev[i], a.fp32, &m, &cc);
        if (lane == 0) {
          amps[d * a.norders + a.fo2] = m;
          corr[d * a.norders + a.fo2] = cc;
        }
        nred += 1;
        phase = kFitDone;
      } else {
        break;
      }
    }
    if (lane == 0) {
      f.ncache = ncache;
      f.phase = phase;
      f.started = started;
      f.angle = angle;
      f.mag = mag;
      f.kov = kov;
      f.n_angle_evals = nang;
      f.n_mag_evals = nmag;
      f.n_reductions = nred;
      f.nreq = w.nreq;
      snreq[d] = w.nreq;
      sdone[d] = phase == kFitDone ? 1 : 0;
    }
    for (int i = lane; i < w.nreq; i += 64) {
      f.req_a[i] = sra[d][i];
      f.req_m[i] = srm[d][i];
    }
  }
  __syncthreads();
  if (active) {
    int c0 = 0, total = 0, done = 0, over = 0;
    for (int j = 0; j < a.ndirs; ++j) {
      if (j < d) c0 += snreq[j];
      total += snreq[j];
      done += sdone[j];
      over |= sover[j];
    }
    if (lane == 0) f.cand0 = c0;
    for (int i = lane; i < snreq[d]; i += 64) {
      // getmodamp's vector at (angle, mag), times order2 - order1 (1093-1094)
      const float ra = sra[d][i], rm = srm[d][i];
      const float kx = rm * dtrig::fcos(ra), ky = rm * dtrig::fsin(ra);
      fs->kc[c0 + i] = make_float2(kx * (float)a.fo2, ky * (float)a.fo2);
      fs->ks[c0 + i] = d;
    }
    if (threadIdx.x == 0) {
      fs->ncand = total;
      fs->ndone = done;
      if (over) fs->overflow = 1;
      if (total > 0) fs->rounds += 1;
    }
  }
}

__global__ void k_fit_kc(const FitState* fs, int ndirs, float mult, float2* kc) {
#pragma clang fp contract(off)
  const int d = threadIdx.x;
  if (d >= ndirs) return;
  const float2 k = fs->d[d].kov;
  kc[d] = make_float2(k.x * mult, k.y * mult);
}

__global__ void k_fit_publish(const FitState* fs, FitArgs a, const float2* amps, const float* corr, FitOut* out,
                              unsigned* flag, unsigned seq) {
  const int t = threadIdx.x;
  if (t < a.ndirs) {
    const FitDir& f = fs->d[t];
    out->k0[t] = f.k0;
    out->kov[t] = f.kov;
    out->n_angle_evals[t] = f.n_angle_evals;
    out->n_mag_evals[t] = f.n_mag_evals;
    out->n_reductions[t] = f.n_reductions;
  }
  for (int i = t; i < a.ndirs * a.norders; i += blockDim.x) {
    out->amps[i] = amps[i];
    out->corr[i] = corr[i];
  }
  if (t == 0) {
    out->ndone = fs->ndone;
    out->rounds = fs->rounds;
    out->overflow = fs->overflow;
  }
  __threadfence_system();  // out is host memory
  __syncthreads();
  if (t == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void check_fit_args(const FitArgs& a) {
  if (a.ndirs < 1 || a.ndirs > kMaxDirs || a.norders < 1 || a.norders > 5 || a.fo2 < 1 || a.fo2 >= a.norders ||
      kFitReq * a.ndirs > kArgCand)
    throw ArgError("device fit: 1..8 directions, 2..5 orders");
}

}  // namespace

void launch_fit_start(FitState* fs, const FitArgs& a, hipStream_t s) {
  check_fit_args(a);
  hipLaunchKernelGGL(k_fit_start, dim3(1), dim3(64), 0, s, fs, a);
  SIMR_LAUNCH_CHECK();
}

void launch_fit_walk(FitState* fs, const FitArgs& a, float2* amps, float* corr, hipStream_t s) {
  check_fit_args(a);
  hipLaunchKernelGGL(k_fit_walk, dim3(1), dim3(64 * a.ndirs), 0, s, fs, a, amps, corr);
  SIMR_LAUNCH_CHECK();
}

void launch_fit_kc(const FitState* fs, int ndirs, float mult, float2* kc, hipStream_t s) {
  hipLaunchKernelGGL(k_fit_kc, dim3(1), dim3(64), 0, s, fs, ndirs, mult, kc);
  SIMR_LAUNCH_CHECK();
}

void launch_fit_publish(const FitState* fs, const FitArgs& a, const float2* amps, const float* corr, FitOut* out,
                        unsigned* flag, unsigned seq, hipStream_t s) {
  check_fit_args(a);
  hipLaunchKernelGGL(k_fit_publish, dim3(1), dim3(64), 0, s, fs, a, amps, corr, out, flag, seq);
  SIMR_LAUNCH_CHECK();
}

}  // namespace simr
