// Host orchestration of one reconstruction context (replaces the reference's
// SIM_Reconstructor internals: setup_common, loadImageData, rescaleDriver,
// processOneVolume and findModulationVectorsAndPhasesForAllDirections,
// src/cudaSirecon/cudaSirecon.cpp:194-494,1139-1321,1354-1435).
#pragma once
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/simr.h"
#include "common.hpp"
#include "stages.hpp"

namespace simr {

// One hipMalloc per context; buffers are bump-allocated from it (replaces the
// reference's per-stage cudaMalloc/cudaFree churn in src/Buffers/GPUBuffer.cpp).
class Arena {
 public:
  void reserve(size_t bytes);
  void* take(size_t bytes);
  void release();
  size_t capacity() const { return cap_; }
  size_t used() const { return off_; }
  ~Arena() { release(); }

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, off_ = 0;
};

struct Layout {
  // sizes derived from the parameters (identical for the footprint query and the context)
  int nx, ny, nz, nxh, ndirs, nphases, norders, nbands, nimg, nDirsOTF;
  long long H, nxy;
  int nx_raw;             // width of the raw sections
  long long nxy_raw;      // nx_raw * ny
  bool resample;          // raw sections are resampled (deskew / crop) into nx-wide planes
  int xdim, ydim, zdim, nasm, zd_ov_max;
  size_t otfsize;
  size_t bytes_total;
  static Layout of(const simr_params& p);
};

class Engine {
 public:
  Engine(const simr_params& p, const float* otf, int device);
  ~Engine();

  void update_params(const simr_params& p);
  void set_raw(const void* src, bool u16, bool on_device);
  void set_flatfield(const float* bg, const float* slope);
  void load_and_rescale(int it);
  void set_time(int it) { cur_t_ = it; }
  void find_modulation();
  void filter_assemble();
  // find_modulation + filter_assemble; a time point that reuses time 0's k0
  // runs as one captured graph (use_graphs)
  void process_volume();
  void filter_dir(int dir);
  void filter_dirs(int d0, int d1, const float2* amps);
  void get_result(float* host);
  const float* result_dev() const { return out_; }
  size_t result_elems() const { return (size_t)L.zdim * L.ydim * L.xdim; }
  simr_fit fit(int dir);
  void set_fit(int dir, const simr_fit& f);
  size_t time0_state_size() const;
  void get_time0_state(void* host, size_t bytes);
  void set_time0_state(const void* host, size_t bytes);
  void get_band(int dir, int band, float* host);
  void set_band(int dir, int band, const float* host);
  void debug_overlaps(int dir, int o1, int o2, float k0x, float k0y, double* P, double* sx, double* sy);
  void overlap_volumes(int dir, int o1, int o2, float k0x, float k0y, float* ov0, float* ov1);
  void sync();
  hipStream_t stream() const { return stream_; }
  const simr_params& params() const { return P; }
  void enable_timing(bool on);
  // Restricts stage timing to the stages of one name ("" = every stage).
  void time_only(const char* name) { timing_only_ = name ? name : ""; }
  int timings(const char** names, float* ms, int cap);

 private:
  struct Geo {
    float dkx, dky, dkz, kzscale, rdistcutoff, lambdaem, lambdaexc, alpha;
  };
  struct Eval {
    double xr, xi, sx, sy;
  };
  // Scratch and state of one direction's k0/modamp fit.  The directions are
  // fitted together: every overlap, cross-correlation and modamp launch
  // covers all of them (one slot per direction), and one read-back serves
  // every direction's line-search step.  The directions' overlap stages and
  // cross-correlations sit one after another (ovx, cc), so a transform pass
  // takes them as one batch; all of it aliases the assembly staging region,
  // which is idle while the fit runs.
  struct FitWork {
    float2* ovx = nullptr;  // [2][nzs_stride][ny][nx]
    double2* P = nullptr;
    double *SX = nullptr, *SY = nullptr;
    float2* cc = nullptr;  // [ny][nx]
    double* sumparts = nullptr;
    int zd = 0;  // z support half-width of the overlaps in ovx
    std::map<std::pair<uint32_t, uint32_t>, Eval> cache;
  };
  struct OvReq {
    int dir, o1, o2;
    float k0x, k0y;
    const float2* k0d = nullptr;  // non-null: the shift is *k0d * kmul, read on the device
    float kmul = 0.f;
  };
  using Walk = std::vector<std::pair<float, float>>;  // (angle, mag) candidates

  void derive();
  Geo geometry() const;
  int zd_overlaps() const;
  int fit_order2() const;
  OtfTab otf_table(int dir, int order) const;
  void wait_fit();
  void wait_flag(unsigned seq);
  OvParams ov_params(int dir, int o1, int o2, float k0x, float k0y) const;
  // builds the requested overlap pairs (each in its direction's slot, or
  // every one in `only` when given) and their z sums
  void make_overlaps(const std::vector<OvReq>& reqs, bool per_elem = false, FitWork* only = nullptr);
  using Cand = std::pair<int, float2>;  // (direction, candidate vector)
  ModampSlots modamp_slots() const;
  void eval_enqueue(const Cand* c, int n, int at);
  void eval_read(int n, int at, Eval* out) const;
  std::vector<std::vector<Eval>> eval_multi(const std::vector<std::pair<int, std::vector<float2>>>& req);
  void modamp_of(const Eval& e, float2* modamp, double* amp2, float* corr) const;
  void prefetch_multi(const std::vector<std::pair<int, Walk>>& req, int o1, int o2);
  bool replay_walk(int dir, bool is_angle, float start, float other, float delta, int o1, int o2, float* out,
                   int* nevals, Walk* need);
  void realspacemodamp_multi(const std::vector<int>& dirs, int o1, int o2, int npend = 0,
                             std::vector<Eval>* pend = nullptr);
  void findk0_all(const std::vector<int>& dirs);
  void fitk0_all(const std::vector<int>& dirs);
  bool fit_on_device();
  void require_finite_k0(int dir, const char* where) const;
  bool graph_eligible() const;
  void build_graph();
  void run_graphed();
  void refresh_fit();
  FilterParams filter_params(int dir, int order) const;
  void assemble();
  int stage_begin(const char* name, hipStream_t s);
  void stage_end(int id, hipStream_t s);

  simr_params P;
  Layout L;
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  Arena arena_;
  bool fit_fp32_ = false;
  float inscale_ = 0;
  std::vector<float> sep_;  // [band][phase]
  std::vector<float> apof_;

  // device buffers
  void* raw_ = nullptr;
  bool raw_u16_ = true, raw_set_ = false, flat_uniform_ = true;
  float *bg_ = nullptr, *slope_ = nullptr, *ratio_ = nullptr, *apofact_ = nullptr;
  double *sums_ = nullptr, *sdp_ = nullptr;
  float2* bands_ = nullptr;
  float2* otf_ = nullptr;
  std::vector<FitWork> fw_;  // one per direction, in the staging region
  FitWork dbg_;              // overlap dumps (its own region: they write the staging region)
  int nzs_stride_ = 1;       // planes reserved per overlap: 2 zd_max + 1
  // batch buffers shared by the directions
  int* slot_ids_ = nullptr;   // 0, 1, ..., kMaxDirs-1 (the graph's candidates)
  double2* part64_ = nullptr;
  float4* part32_ = nullptr;
  float* pv_ = nullptr;
  int* pi_ = nullptr;
  // Coherent pinned host memory: the modamp sums, the cross-correlation
  // peaks and the round's completion flag are written there by the kernels
  // directly (the candidates travel in the kernel arguments), so a
  // line-search round has no copy and one wait.  Host and device views of
  // the same bytes.
  char* hzc_ = nullptr;
  char* dzc_ = nullptr;
  unsigned fit_seq_ = 0;  // the last completion flag value queued
  // device-driven fit (fit_on_device): its state, in the fit scratch; the
  // (evaluate, walk) rounds queued per fit, adapted to the rounds the last
  // fit needed
  FitState* fitst_ = nullptr;
  int fit_rounds_ = 4;
  bool device_fit_ = true;
  bool fused_ov_ = true;  // two-pass overlap construction (launch_overlaps_fused_cols/_rows)
  float2 *stageZ_ = nullptr, *stageY_ = nullptr;
  float* resampled_ = nullptr;  // [nimg][ny][nx] resampled raw planes (aliases the staging region)
  float* out_ = nullptr;

  size_t stage_bytes_ = 0;  // bytes of the assembly staging region

  // reconstruction state
  int cur_t_ = 0;
  bool preprocessed_ = false, fitted_ = false, processed_ = false;
  std::vector<float2> k0_, k0_time0_, k0guess_;
  std::vector<std::vector<float2>> amp_;
  std::vector<std::vector<float>> corr_;
  std::vector<simr_fit> stats_;

  // device-side modamps of the graphed time points
  float2* kc_dev_ = nullptr;    // [ndirs][norders] candidate vectors k0 * order
  float2* amps_dev_ = nullptr;  // [ndirs][norders] modamps (order 0: 1)
  float* corr_dev_ = nullptr;   // [ndirs][norders] correlation coefficients
  void* fit_host_ = nullptr;    // pinned read-back of amps_dev_ and corr_dev_
  bool fit_pending_ = false;    // fit_host_ holds a read-back not yet moved into amp_ / corr_
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<float2> graph_k0_;  // k0 the graph was captured with
  int graph_version_ = -1, param_version_ = 0;

  // timing
  bool timing_ = false;
  std::string timing_only_;
  int cur_stage_ = -1;  // open stage on the context stream
  std::mutex ev_mu_;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> events_;
  std::vector<hipEvent_t> ev_pool_;  // timing events not in use (created once, reused)
  std::vector<std::pair<std::string, float>> last_timings_;
};

}  // namespace simr
