// stepsim — FluidSystem::stepSimulate (fluids/FluidSystem.cpp:99-120) as a
// C++ host program over the C++ mirror of the reference's Simulator
// (include/pbf_simulator.hpp): per frame the box of the sweeping wall is set
// with setLim (fluids/FluidSystem.cpp:104-110), then one step with tic-toc
// buffers (fluids/FluidSystem.cpp:112-117).
//
//   stepsim IN OUT
//
// IN (little-endian): int32 n, int32 nframes, pbf_params, pbf_sweep, then
// pos (n x 4 float32), vel (n x 4 float32), iid (n uint32).
// OUT: the newest pos, vel (n x 4 float32) and iid (n uint32) after nframes.
// Exit status: 0 done, 1 library error, 2 usage or file error, 3 HIP error.
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pbf_simulator.hpp"

namespace {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "stepsim: %s: %s\n", what, hipGetErrorString(e));
    std::exit(3);
  }
}

bool read_all(std::FILE* f, void* dst, size_t bytes) {
  return std::fread(dst, 1, bytes, f) == bytes;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: stepsim IN OUT\n");
    return 2;
  }
  std::FILE* in = std::fopen(argv[1], "rb");
  if (!in) {
    std::fprintf(stderr, "stepsim: cannot open %s\n", argv[1]);
    return 2;
  }
  int32_t hdr[2];
  pbf_params p;
  pbf_sweep sw;
  if (!read_all(in, hdr, sizeof hdr) || !read_all(in, &p, sizeof p) ||
      !read_all(in, &sw, sizeof sw) || hdr[0] <= 0 || hdr[1] < 0) {
    std::fprintf(stderr, "stepsim: bad header in %s\n", argv[1]);
    std::fclose(in);
    return 2;
  }
  const int n = hdr[0], nframes = hdr[1];
  const size_t vec_bytes = sizeof(float) * 4 * (size_t)n, id_bytes = sizeof(uint32_t) * (size_t)n;
  std::vector<float> pos(4 * (size_t)n), vel(4 * (size_t)n);
  std::vector<uint32_t> iid((size_t)n);
  const bool ok = read_all(in, pos.data(), vec_bytes) && read_all(in, vel.data(), vec_bytes) &&
                  read_all(in, iid.data(), id_bytes);
  std::fclose(in);
  if (!ok) {
    std::fprintf(stderr, "stepsim: %s is shorter than its header says\n", argv[1]);
    return 2;
  }

  float *d_pos, *d_npos, *d_vel, *d_nvel;
  uint32_t* d_iid;
  hip_check(hipMalloc(&d_pos, vec_bytes), "hipMalloc");
  hip_check(hipMalloc(&d_npos, vec_bytes), "hipMalloc");
  hip_check(hipMalloc(&d_vel, vec_bytes), "hipMalloc");
  hip_check(hipMalloc(&d_nvel, vec_bytes), "hipMalloc");
  hip_check(hipMalloc(&d_iid, id_bytes), "hipMalloc");
  hip_check(hipMemcpy(d_pos, pos.data(), vec_bytes, hipMemcpyHostToDevice), "hipMemcpy");
  hip_check(hipMemcpy(d_vel, vel.data(), vec_bytes, hipMemcpyHostToDevice), "hipMemcpy");
  hip_check(hipMemcpy(d_iid, iid.data(), id_bytes, hipMemcpyHostToDevice), "hipMemcpy");

  bool tictoc = false;
  try {
    pbf_vec3 ul, ll;
    pbf::check(pbf_sweep_box(&sw, 0, &ul, &ll), nullptr);
    pbf::Simulator sim(p, ul, ll, n, 0, PBF_LAYOUT_FLOAT4);
    for (int frame = 0; frame < nframes; ++frame) {
      pbf::check(pbf_sweep_box(&sw, frame, &ul, &ll), nullptr);
      sim.setLim(ul, ll);
      if (!tictoc)
        sim.step(d_pos, d_npos, d_vel, d_nvel, d_iid, n);
      else
        sim.step(d_npos, d_pos, d_nvel, d_vel, d_iid, n);
      tictoc = !tictoc;
    }
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  } catch (const pbf::Error& e) {
    std::fprintf(stderr, "stepsim: %s\n", e.what());
    return 1;
  }

  hip_check(hipMemcpy(pos.data(), tictoc ? d_npos : d_pos, vec_bytes, hipMemcpyDeviceToHost),
            "hipMemcpy");
  hip_check(hipMemcpy(vel.data(), tictoc ? d_nvel : d_vel, vec_bytes, hipMemcpyDeviceToHost),
            "hipMemcpy");
  hip_check(hipMemcpy(iid.data(), d_iid, id_bytes, hipMemcpyDeviceToHost), "hipMemcpy");
  for (void* d : {(void*)d_pos, (void*)d_npos, (void*)d_vel, (void*)d_nvel, (void*)d_iid})
    hip_check(hipFree(d), "hipFree");

  std::FILE* out = std::fopen(argv[2], "wb");
  if (!out) {
    std::fprintf(stderr, "stepsim: cannot open %s\n", argv[2]);
    return 2;
  }
  const bool wrote = std::fwrite(pos.data(), 1, vec_bytes, out) == vec_bytes &&
                     std::fwrite(vel.data(), 1, vec_bytes, out) == vec_bytes &&
                     std::fwrite(iid.data(), 1, id_bytes, out) == id_bytes;
  if (std::fclose(out) != 0 || !wrote) {
    std::fprintf(stderr, "stepsim: cannot write %s\n", argv[2]);
    return 2;
  }
  return 0;
}
