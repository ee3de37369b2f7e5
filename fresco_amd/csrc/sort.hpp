// Stable radix sort of uint32 (key, value) pairs on the device, built from
// kernel launches only (no memory fills), so it can be captured in a graph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace fresco_hip {

// Records carried along by the sort: element i of the output gets
// rec_in[v] (rec_words 32-bit words, a multiple of 4) and cell_in[v], v the
// output value (the original index).  Moved in the last digit pass only.
struct SortPayload {
  const void* rec_in = nullptr;
  void* rec_out = nullptr;
  int rec_words = 4;
  const uint32_t* cell_in = nullptr;
  uint32_t* cell_out = nullptr;
};

class BinSort {
 public:
  // n pairs; keys are below 2^key_bits
  void init(int64_t n, int key_bits);
  // the same, with the count tables (and the ping-pong buffers of
  // multi-pass sorts) carved from ws instead of allocated
  void init(int64_t n, int key_bits, Carve& ws);
  // bytes init(n, key_bits, ws) takes from ws
  static size_t workspace_bytes(int64_t n, int key_bits);
  // keys / vals -> keys_out / vals_out (stable; inputs are not modified);
  // with a payload, its records are placed in the output order too
  void sort(const uint32_t* keys, const uint32_t* vals, uint32_t* keys_out, uint32_t* vals_out, hipStream_t s,
            const SortPayload* payload = nullptr);
  // Keys seg * nbuck + digit whose input is already grouped by seg in
  // segments [starts[i], starts[i+1]) (starts: host, nseg + 1 entries):
  // one stable pass on the digit inside every segment gives the same output
  // as sort().  Call once after init; sort() then takes this path when its
  // staging fits 64 KB of LDS (else the passes over the combined keys).
  void set_segments(const std::vector<int64_t>& starts, int nbuck);

 private:
  int64_t n_ = 0;
  int passes_ = 1;
  int dbits_ = 1;  // digit width, at most 11 bits
  int nblk_ = 1;
  int nseg_ = 1;
  DeviceBuffer<uint32_t> cnt_, tops_, ktmp_, vtmp_;
  // segmented mode
  bool seg_ = false;
  int nseg_tiles_ = 0, nsegs_ = 0, seg_buck_ = 0;
  DeviceBuffer<int64_t> tile_beg_, tile_end_, seg_start_;
  DeviceBuffer<int> seg_tile0_;
  DeviceBuffer<uint32_t> seg_cnt_;
};

}  // namespace fresco_hip
