// program.hpp -- a circuit bound to a device state: the launch list the executor
// replays, the device-resident matrices it reads, and the optional captured graph.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "engine.hpp"

namespace qsvh {

struct Program {
  State* st = nullptr;
  std::vector<Launch> launches;
  void* arena = nullptr;  // device matrices (state precision, interleaved re/im)
  size_t arena_bytes = 0;
  bool use_graph = false;
  hipGraphExec_t graph = nullptr;
  int precision = 0;
  void* mfma = nullptr;      // block matrices of the MFMA group launches
  std::vector<void*> owned;  // further device buffers freed with the program
  // Tile passes: Launch::kTile launches index `tile_passes` through arena_off.
  std::vector<TilePassArg> tile_passes;
  TileOp* tile_ops = nullptr;
  TileDiag* tile_diags = nullptr;
  void* tile_pool = nullptr;
  // Original op index of each applied gate, in execution order (tile passes reorder
  // gates that act on disjoint qubits).
  std::vector<uint64_t> order;
};

int set_err(Ctx* ctx, int code, const std::string& msg);
int hip_err(Ctx* ctx, hipError_t e, const char* where);

// Passes that rewrite the plain launch list (plan.cpp).
int plan_tiles(Program* p, int tile_qubits);
int plan_mfma(Program* p);
void release_extra(Program* p);

}  // namespace qsvh
