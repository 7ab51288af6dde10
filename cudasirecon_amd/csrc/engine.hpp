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
  // Scratch and state of one direction's k0/modamp fit.  Directions are
  // fitted concurrently, each on its own stream and host thread; direction
  // 0 uses the context stream, the others alias the assembly staging
  // buffers, which are idle while the fit runs.
  struct FitWork {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    float2* ovx = nullptr;
    double2* P = nullptr;
    double *SX = nullptr, *SY = nullptr;
    float2* cc = nullptr;
    float* pv = nullptr;
    int* pi = nullptr;
    float* peakvals = nullptr;
    int* peakidx = nullptr;
    float2* kcand = nullptr;
    double2* part64 = nullptr;
    float4* part32 = nullptr;
    double* sumparts = nullptr;
    void* hstage = nullptr;  // pinned host staging
    hipEvent_t done = nullptr;  // read-back completion (polled by wait_fit)
    int zd = 0;  // z support half-width of the overlaps in ovx
    bool ov_totals_valid = false;
    double ov_sx = 0, ov_sy = 0;
    std::map<std::pair<uint32_t, uint32_t>, Eval> cache;
    int cache_o1 = -1, cache_o2 = -1;
  };

  void derive();
  Geo geometry() const;
  int zd_overlaps() const;
  int fit_order2() const;
  OtfRA otf_table(int dir, int order) const;
  void carve_fit(FitWork& w, char*& at, bool count_only, size_t& bytes) const;
  void wait_fit(FitWork& w);
  void make_overlaps(FitWork& w, int dir, int o1, int o2, float k0x, float k0y, bool per_elem = false);
  std::vector<Eval> eval_batch(FitWork& w, const std::vector<float2>& kvecs);
  void modamp_of(const Eval& e, float2* modamp, double* amp2, float* corr) const;
  double getmodamp(FitWork& w, int dir, float angle, float mag, int o1, int o2, float2* modamp, float* corr);
  double realspacemodamp(FitWork& w, int dir, int o1, int o2, float2 kv, bool redo, float2* modamp, float* corr);
  void prefetch(FitWork& w, int o1, int o2, const std::vector<std::pair<float, float>>& am);
  void findk0(FitWork& w, int dir);
  void fitk0andmodamps(FitWork& w, int dir);
  void modulation(FitWork& w, int dir);
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
  std::vector<FitWork> fw_;  // one per direction
  float2 *stageZ_ = nullptr, *stageY_ = nullptr;
  float* resampled_ = nullptr;  // [nimg][ny][nx] resampled raw planes (aliases the staging region)
  float* out_ = nullptr;

  size_t hstage_bytes_ = 0;
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
  hipEvent_t fork_ev_ = nullptr;
  std::vector<hipEvent_t> join_ev_;

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
