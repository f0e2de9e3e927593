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

  pbf_solver* handle() { return s_; }

 private:
  pbf_solver* s_ = nullptr;
};

}  // namespace pbf

#endif  // PBF_SIMULATOR_HPP_
