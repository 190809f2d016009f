// Fused assembly form of a HessianAssembly: one work tile per compact patch of owner nodes.
//
// The waves form of an element-sum Hessian writes every element's Hessian block to HBM and
// reads it back in a second launch through a per-contribution gather index; on a cloth that
// is five times the algorithmic traffic.  The fused form removes both: a tile owns a set of
// nodes (every output column of those nodes), evaluates every element touching them into LDS,
// and sums each owned output from LDS into its final CSC position.  Elements on the boundary
// of two tiles are evaluated by both (the recompute factor is reported).
//
// Node layout: instance variables are numNodes x dofs, column-major (variable = node + dof *
// numNodes), the layout of an n x 3 position matrix; an element's variables are dofs of its
// nodes.  A tile stages the coordinates of its nodes (every node of its elements, each once)
// as x[base + node + dof * numNodes] from a node list, and its elements reach them through
// u16 local node slots.
//
// Tile pool (u32 words, read by the generated kernel and checked by validateFused):
//   pool[0]              number of tiles
//   pool[1 + t]          offset of tile t's header, tiles listed in launch order
//   header H (offsets marked rel are relative to H):
//     H[0] input base of the instance     H[1] numNodes of the instance
//     H[2] m, owned outputs               H[3] 64-output chunks
//     H[4] rel runs: (result position, first output) pairs, one sentinel (first = m)
//     H[5] rel counts: u8 per output, element contributions (bits 0-6) and whether the
//          output adds a constant (bit 7)
//     H[6] rel slots: u16 LDS byte offsets of element values (8 x value index), contributions
//          of every output in order
//     H[7] rel chunks: (first slot, run of the first output, first constant, run-start mask
//          low and high: bit j set when output 64 ch + j, j >= 1, starts a run) per 64-output
//          chunk, FUSED_CHUNK_WORDS words each
//     H[8] output constants in c[], one per output with bit 7 set, in output order, or
//          0xffffffff
//     H[9] result base of the instance
//     H[10] words from runs to the end of chunks (the recipe, contiguous)
//     H[11] stored output constants (bits 0-27) and their period (bits 28-31): 0, or D when
//          the tile's outputs form D equal dof blocks (columns b + d n) holding the same
//          constants in the same order, of which one block's are stored
//     H[12] the tile's element image in the value buffer (split form: first value, in doubles)
//     H[13] nn, nodes of the tile's elements    H[14] rel tile node list (nn u32, sorted)
//     H[15 + 3 t], H[16 + 3 t], H[17 + 3 t]: template t's element count, rel slot table
//       (u16 local node slots, numNodes_t columns of count, packed two per word), parameter
//       block in c[] (numParams_t columns of count)
//   LDS of a tile, in doubles: element values (template t's outputs at V_t + i * stride_t
//   + slot, V_t = sum over t' < t of count_t' * stride_t', stride = valueStride), the zero
//   V[nvals], then the staged region: parameters (template t: parameter k at P_t + k *
//   count_t + i), node coordinates (dof d of local node s at X + d * nn + s, stagedDofs
//   dofs), slot tables (template t's words at W_t), then the recipe (u32 words copied from
//   H[4]), then the output constants.
//
// Block sums (FusedProgram::sums == FusedSums::Blocks).  When every element adds a full D x D
// block to every pair of its nodes (D = blockDofs: all dofs of a node are differentiated) and
// constants are nodal blocks M (x) I_D, the Hessian is made of node blocks: block (a, b) of
// column node b is either full (a shares an element with b; a == b is the diagonal block) or
// holds only its diagonal (a is coupled to b by a constant alone, a "wing").  A lane then sums
// a whole block from one list of (element value base, transposed) entries, since the D x D
// entries of one element's block sit at consecutive value slots (the template's outputs are
// laid out block by block: pair (p, q), p < q, of the element's nodes as D x D row-major
// values, pair (p, p) as its D (D + 1) / 2 upper-triangle values).  Column (b, db) of the
// output holds, for da = 0..D-1, the nodes E(b) (sharing an element with b, b included) and,
// when da == db, the wings K(b), each segment sorted by node; so entry (a, da; b, db) of a
// block sits at tile output
//   cs[db] + da * nE + (db < da ? nK : 0) + (da == db ? rank of a in E(b) u K(b) : rank in E(b)).
// Lanes keep their block sums in registers (a local array beyond 8 blocks per thread), write
// them after a barrier to the tile's output slots in LDS (over the element values, which are
// dead by then), and a last pass stores the slots in CSC order, a pair of consecutive slots per
// lane (16 bytes: one LDS read, one store; a wave stores whole lines).  Slots follow the
// tile's runs; every run starts at an even slot (a padding slot before it when needed, and
// one at the end when the count is odd), so a pair never spans two runs; cs[] above counts
// slots.  Waves take the 64-block chunks in snake order (every second pass reversed).  Header
// words that differ from the form above:
//     H[2] m2, output slots of the tile (outputs and padding, even)   H[3] nb, blocks
//     H[4] rel recipe start: output chunk records per 128 slots (run of the first pair, pairs
//          after it where a run starts, pairs whose second slot is padding: 5 words), then
//          block chunk records (list base in words | longest list << 24, first constant |
//          halved << 31: 2 words per 64 block records), then block records (1 word per lane of
//          a chunk: owned node index bits 0-13, kind bits 14-15 (0 diagonal, 1 full, 2 wing,
//          3 none: 0xc000), rank in E(b) bits 16-22, rank in E(b) u K(b) bits 23-29, has a
//          constant bit 30)
//     H[5] rel runs: (result position, first slot) pairs, one sentinel (first = m2)
//     H[6] rel lists   H[7] rel node records (2 words per owned node: cs[0] | cs[1] << 16,
//          cs[2] | nE << 16 | nK << 24)
//     H[8] block constants in c[] (one per block that has one, in block order) or 0xffffffff
//     H[10] words from the recipe start to the end of the lists   H[11] stored constants
//   Lists: a chunk of 64 blocks reads (longest + 1) / 2 u32 words per lane, word j of lane l at
//   list base + 64 j + l, each two u16 entries (value index bits 0-14, transposed bit 15);
//   entries past a block's list point at D x D zeros after the element values (V[nvals ..]).
//   Halved chunks come first when the diagonal blocks' lists are at least 4 long and twice the
//   full blocks' longest: 32 diagonal blocks each, lane l summing the first (len + 1) / 2
//   entries of block l and lane l + 32 the rest (its record: lane l's without the constant);
//   the halves meet in lane l through a cross-lane add, and lane l + 32 writes nothing.
//   LDS: values and their zeros, or the tile's slots when more (max(nvals + D^2, m2) doubles),
//   then the staged region and the recipe; a lane reads its block's constant from c[] (one
//   load, issued before the chunk's sums), so constants take no LDS.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "Program.hpp"
#include "Symbolic.hpp"

namespace Sym {

enum class FusedSchedule : int {
    Static = 0,       // one workgroup per tile, tiles in build order
    Queue = 1,        // persistent workgroups pull tiles from an atomic counter, build order
    QueueLargest = 2, // persistent, largest tiles first
    Interleaved = 3,  // persistent, tiles of different instances interleaved round-robin
    Strided = 4       // persistent without a counter: workgroup b takes tiles b, b + grid, ...
};

struct FusedOptions {
    size_t ldsBytes = 53760;          // LDS bytes per workgroup, all regions (3 fit in 160 KiB)
    size_t maxOutputs = 1u << 16;     // owned outputs per tile
    int threadsPerBlock = 256;
    FusedSchedule schedule = FusedSchedule::Queue;
    int outputGroups = 5;             // work items per element (outputs split between them)
    int wavesPerEU = 0;               // minimum waves per SIMD asked of the compiler (0 = none)
    bool profile = false;             // record wall-clock stamps per tile (see tileProfileWords)
    bool shareReciprocals = true;     // one 1/d per repeated denominator d (x/d -> x*(1/d))
    bool foldCoefficients = true;     // carry constant factors through the statements
    // split form: one launch evaluates every tile's elements into its image in HBM (no LDS for
    // values, high occupancy), a second launch stages images and recipes in LDS and sums
    bool split = false;
    // stage tiles with direct global-to-LDS loads (no register round trip, no wait per load);
    // false: through registers
    bool ldsDma = true;
    // element tiles: sum node blocks when the Hessian is block-structured (FusedSums::Blocks);
    // false: one contribution list per output
    bool blockSums = true;
};

enum class FusedSums : int { Outputs = 0, Blocks = 1 };

// `work` starts with two work-queue counters (next tile, finished workgroups) per launch of an
// evaluation (the split form has two launches).  Profiled launches record after them, per tile
// in launch order, four 64-bit wall-clock stamps (tile start, staging done, element evaluation
// done, outputs written; in the split form the last stamp comes from the sum launch).
constexpr size_t WORK_COUNTER_WORDS = 4;
inline size_t tileProfileWords(size_t numTiles) { return WORK_COUNTER_WORDS + 8 * numTiles; }

struct FusedTemplate {
    Symbolic code;                 // ASSIGN(VAR(s,-2), ...) over VAR(a,-3) arguments, VAR(q,-4) parameters
    int numNodes = 0;              // nodes per element
    std::vector<int> argNode, argDof;  // per gathered argument
    int numParams = 0;             // parameter columns (non-uniform constants)
    int numOutputs = 0;            // values per element
    // LDS doubles between consecutive elements' values (numOutputs rounded up to odd, so the
    // 64 lanes of a wave storing one output of 64 consecutive elements hit distinct banks)
    int valueStride = 0;
};

struct FusedProgram {
    int kind = 0;  // 0: element sums (HessianAssembly), 1: product sums (A^T A + A)
    std::vector<FusedTemplate> templates;
    std::vector<uint32_t> pool;
    std::vector<double> consts;
    // per tile in build order: header offset, instance, index within the instance, cost
    // estimate (outputs + element values); the launch order in pool[1..] is derived from it
    std::vector<uint32_t> tileHeader, tileInstance, tileIndex;
    std::vector<float> tileCost;
    size_t numInputs = 0, numResults = 0;
    std::vector<size_t> outStart, outLength;
    size_t numTiles = 0, ldsBytes = 0, elementLdsBytes = 0;  // per workgroup: all LDS, element part
    size_t stagedLdsBytes = 0;  // split form, element launch: staged arguments and parameters
    size_t imageDoubles = 0;    // split form: element images of all tiles (after the inputs)
    // element tiles: copies of the headers in launch order at pool[headerTable + t * headerStride]
    // (headerTable = 1 + numTiles, right after the launch order) with the relative offsets
    // (H[4..7], node lists) replaced by absolute pool indices, each followed at headerNodes by a
    // copy of the tile's node list (H[14] points there), so a workgroup loads its header and its
    // node numbers in one round of loads
    size_t headerTable = 0, headerStride = 0, headerNodes = 0;
    int stagedDofs = 0;  // dofs staged per tile node: 1 + the largest dof an element reads
    int maxPairs = 0;    // element tiles: the most slot pairs (even-padded contributions / 2) of an output
    FusedSums sums = FusedSums::Outputs;
    int blockDofs = 0;      // block sums: D, dofs per node block
    size_t maxBlocks = 0;   // block sums: the most blocks of a tile
    size_t elementEvaluations = 0, numElements = 0, numContributions = 0;
    double recompute() const { return numElements ? (double)elementEvaluations / (double)numElements : 0.0; }
    std::string summary() const;
};

// Per tile in launch order, TILE_STATS floats: owned outputs, element contributions, elements,
// staged doubles, element LDS values, recipe words, output constants, runs.
constexpr int TILE_STATS = 8;
constexpr int FUSED_HEADER = 15;  // fixed header words of an element tile
// Header words holding pool offsets (relative in a tile's header, absolute in the table).
inline bool fusedHeaderOffset(size_t w) {
    return (w >= 4 && w <= 7) || w == 14 || (w >= (size_t)FUSED_HEADER && (w - FUSED_HEADER) % 3 == 1);
}
// Doubles of a tile's staged region (parameters, node coordinates, slot tables) from its header.
size_t fusedStagedDoubles(const FusedProgram& fp, const uint32_t* H);
constexpr int FUSED_CHUNK_WORDS = 5;  // words per 64-output chunk record of an element tile
// Element tiles whose outputs have at most this many slot pairs sum with every pair's reads
// issued at once (pairs past a lane's list replaced by the zero slot); longer lists loop.  A
// lane past a tile's last output reads that many words from the end of the slot table, inside
// the chunk records that follow it (at least FUSED_CHUNK_WORDS words).
constexpr int FUSED_UNROLLED_PAIRS = FUSED_CHUNK_WORDS;
constexpr size_t FUSED_MAX_LDS_BYTES = 160 * 1024;  // LDS of one gfx950 CU
constexpr size_t FUSED_MAX_BLOCKS = 4096;           // block sums: node blocks of one tile
std::vector<float> tileStats(const FusedProgram& fp);

// Write the tile table pool[1..numTiles] in the order of `schedule`.
void orderTiles(FusedProgram& fp, FusedSchedule schedule);

// Product-sum form of B = A^T A + A for a sparse A (CSC, rows sorted and unique per column):
// B(i, j) = A(i, j) + sum_k A(k, i) A(k, j), the structural product (no cancellation drops an
// entry), as the reference's Eigen expression AS.transpose() * AS + AS
// (examples/Tutorial/main.cpp:43) evaluates symbolically.  A tile owns consecutive output
// columns; it stages the values of A its products read (the rows of A that meet its columns)
// in LDS, and each output sums its products from LDS.
// Tile header: H[0], H[1] unused, H[2] .. H[11] as in the element form (H[8] = none,
// H[11] = 0), H[12] staged values, H[13] rel staged positions (one u32 position in A's values
// per staged value, in LDS order).  Counts: bits 0-6 products, bit 7 whether the output adds
// A(i, j); slots: [A(i, j)] then two LDS indices per product.
FusedProgram buildProductTiles(int n, const int* colptr, const int* rowidx, const FusedOptions& fo,
                               std::vector<int>& outer, std::vector<int>& inner);

// Walks the pool as the kernel does and checks every index against its buffer; throws
// std::logic_error on the first violation.
void validateFused(const FusedProgram& fp);

// What the emitted element code does per element, summed over the templates and their output
// groups: FP64 operations, statements, and the most values any one work item keeps live.
struct FusedCodeStats {
    size_t adds = 0, muls = 0, divs = 0, sqrts = 0, statements = 0;
    int maxLive = 0;
};

// Generated source: one __global__ symlib_fused, per template a staging function and one
// device function per output group.
std::string emitHipFused(const FusedProgram& fp, const FusedOptions& opt, std::vector<LaunchDesc>& launches,
                         FusedCodeStats* stats = nullptr);

}  // namespace Sym
