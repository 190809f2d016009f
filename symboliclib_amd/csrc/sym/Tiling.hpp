// Tile-fused execution of a Program: the LDS-staged work tiles of the MI355X back-end.
//
// In the wave form (Program + emitHip) every intermediate value (a per-face Hessian entry, a
// per-corner cotangent weight, ...) is written to HBM by one launch and gathered back by the
// next.  The tiled form removes that round trip: the user outputs are partitioned into tiles,
// and the workgroup of a tile recomputes exactly the intermediate instances its outputs need
// (the closure) into LDS, then evaluates its outputs from LDS.  Instances needed by several
// tiles are recomputed by each of them; arithmetic is cheap next to HBM traffic.
//
// Tiles are grown greedily: starting from the first unassigned output (in output order),
// the outputs that add the fewest new intermediate instances are added first, until the
// LDS budget or the output cap is reached.  On mesh-like programs this grows compact patches,
// so the recomputed halo stays small.
//
// Per tile the device reads one u32 table (tile pool): for every kernel its instance count,
// data offset and LDS offset, then per kernel and instance the gather operands (an input
// position, or LDS_FLAG | LDS index for an intermediate), the global instance id (kernels with
// constant tables) and the output positions (kernels writing user outputs).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "Program.hpp"

namespace Sym {

constexpr uint32_t LDS_FLAG = 0x80000000u;

struct TileOptions {
    size_t ldsBytes = 64 * 1024;   // LDS budget of one tile
    size_t maxOutputs = 8192;      // user-output instances per tile
    int threadsPerBlock = 256;
};

struct TiledProgram {
    bool valid = false;
    std::string reason;            // why tiling was not possible
    size_t numTiles = 0;
    size_t ldsBytes = 0;           // largest LDS use of a tile
    std::vector<uint32_t> pool;    // [tile header offsets (numTiles + 1) | headers | data]
    // per kernel: gather column kinds (0 input, 1 LDS, 2 mixed) at colOffset[k] ...
    std::vector<uint8_t> colKind;
    std::vector<size_t> colOffset;
    std::vector<char> kernelUsed;  // intermediate kernels that no output needs are skipped
    size_t uniqueIntermediates = 0, tiledIntermediates = 0;  // recomputation = tiled / unique
    size_t numOutputInstances = 0;
};

TiledProgram buildTiles(const Program& prog, const TileOptions& opt);

// Bounds check of every table entry the tiled kernel dereferences (throws std::logic_error).
void validateTiles(const Program& prog, const TiledProgram& tp);

// HIP source of the tiled form: one extern "C" kernel "symlib_tiled", one workgroup per tile.
std::string emitHipTiled(const Program& prog, const TiledProgram& tp, const EmitOptions& opt,
                         std::vector<LaunchDesc>& launches);

}  // namespace Sym
