// pbf_simulator.hpp — C++ mirror of the reference's Simulator class over the
// C ABI of pbf.h (header only).
//
// Same names and argument meaning as naeioi/PBF-CUDA's Simulator
// (fluids/Simulator.h:7-61): construct with parameters and box, step() one
// frame, loadParams()/saveParams(), setLim().  Differences: buffers are
// device pointers (the reference takes GL buffer names and maps them every
// step, fluids/Simulator.cpp:16-36), parameters are passed in rather than
// read from the GUIParams singleton, and failures throw pbf::Error instead of
// exiting the process (checkCudaErrors).
#ifndef PBF_SIMULATOR_HPP_
#define PBF_SIMULATOR_HPP_

#include <stdexcept>
#include <string>

#include "pbf.h"

namespace pbf {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg)
      : std::runtime_error(std::string(pbf_status_string(code)) + ": " + msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const pbf_solver* s) {
  if (rc != PBF_OK) throw Error(rc, pbf_last_error(s));
}

// Simulator(params, ulim, llim) (fluids/Simulator.h:10); capacity replaces
// MAX_PARTICLE_NUM (fluids/helper.h:10).
class Simulator {
 public:
  Simulator(const pbf_params& params, pbf_vec3 ulim, pbf_vec3 llim, int capacity,
            long long max_cells = 0, int layout = PBF_LAYOUT_FLOAT4) {
    check(pbf_create(&params, ulim, llim, capacity, max_cells, layout, &s_), nullptr);
  }
  ~Simulator() { pbf_destroy(s_); }
  Simulator(const Simulator&) = delete;
  Simulator& operator=(const Simulator&) = delete;

  // Simulator::step (fluids/Simulator.h:37); stream is a hipStream_t.
  void step(float* pos, float* npos, float* vel, float* nvel, uint32_t* iid, int nparticle,
            void* stream = nullptr) {
    check(pbf_step(s_, pos, npos, vel, nvel, iid, nparticle, stream), s_);
  }
  // Simulator::loadParams (fluids/Simulator.cpp:101-115), explicit source.
  void loadParams(const pbf_params& p) { check(pbf_set_params(s_, &p), s_); }
  // Simulator::saveParams (fluids/Simulator.cpp:117-130), returned.
  pbf_params saveParams() const {
    pbf_params p;
    check(pbf_get_params(s_, &p), s_);
    return p;
  }
  // Simulator::setLim (fluids/Simulator.cpp:132-136).
  void setLim(const pbf_vec3& ulim, const pbf_vec3& llim) {
    check(pbf_set_lim(s_, ulim, llim), s_);
  }

  // Additions without a reference counterpart: HIP graph replay, launch
  // shapes and the autotuner (pbf.h).
  void setGraph(bool enable) { check(pbf_set_graph(s_, enable ? 1 : 0), s_); }
  void setShapes(int light_block, int chunk_blocks) {
    const pbf_shapes sh{light_block, chunk_blocks};
    check(pbf_set_shapes(s_, &sh), s_);
  }
  // 0 = adaptive (default); results never depend on it.
  void setHeavyThreshold(int threshold) { check(pbf_set_heavy_threshold(s_, threshold), s_); }
  pbf_shapes autotune(const float* pos, const float* vel, const uint32_t* iid, int n, int reps = 3,
                      void* stream = nullptr, double* best_ms = nullptr) {
    pbf_shapes best;
    check(pbf_autotune(s_, pos, vel, iid, n, reps, stream, &best, best_ms), s_);
    return best;
  }

  pbf_solver* handle() { return s_; }

 protected:
  Simulator() = default;
  pbf_solver* s_ = nullptr;
};

// Many boxes stepped together (pbf_create_batched): scene k owns indices
// [off_k, off_k + counts[k]) of every array.
class BatchedSimulator : public Simulator {
 public:
  BatchedSimulator(const pbf_params& params, int nscenes, const pbf_vec3* ulim,
                   const pbf_vec3* llim, const int* counts, long long max_cells = 0,
                   int layout = PBF_LAYOUT_FLOAT4) {
    check(pbf_create_batched(&params, nscenes, ulim, llim, counts, max_cells, layout, &s_),
          nullptr);
  }
  void stepBatched(float* pos, float* npos, float* vel, float* nvel, uint32_t* iid,
                   void* stream = nullptr) {
    check(pbf_step_batched(s_, pos, npos, vel, nvel, iid, stream), s_);
  }
};

// Checkpointed multi-frame run of a single-scene Simulator and its adjoint.
class Tape {
 public:
  // flags: PBF_TAPE_STORE keeps the iterations instead of recomputing them.
  Tape(Simulator& sim, int n, int nframes, int flags = 0) : sim_(sim) {
    check(pbf_tape_create_ex(sim.handle(), n, nframes, flags, &t_), sim.handle());
  }
  ~Tape() { pbf_tape_destroy(t_); }
  Tape(const Tape&) = delete;
  Tape& operator=(const Tape&) = delete;
  void forward(const float* pos, const float* vel, const uint32_t* iid, void* stream = nullptr,
               int first_frame = 0, const pbf_sweep* sweep = nullptr) {
    check(pbf_tape_forward(sim_.handle(), t_, pos, vel, iid, first_frame, sweep, stream),
          sim_.handle());
  }
  // Returns dL/dpho0; gpos0/gvel0 receive dL/d(initial positions, velocities).
  double backward(const float* gpos, const float* gvel, float* gpos0, float* gvel0,
                  void* stream = nullptr) {
    double g = 0.0;
    check(pbf_tape_backward(sim_.handle(), t_, gpos, gvel, gpos0, gvel0, &g, stream),
          sim_.handle());
    return g;
  }
  pbf_tape* handle() { return t_; }

 private:
  Simulator& sim_;
  pbf_tape* t_ = nullptr;
};

}  // namespace pbf

#endif  // PBF_SIMULATOR_HPP_
