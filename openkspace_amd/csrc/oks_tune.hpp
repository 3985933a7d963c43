// oks_tune.hpp -- launch-shape table written by the project's autotuner
// (python -m openkspace_amd.autotune) and read at context creation.
//
// The table maps (op, shape key) to launch-shape options for one GPU
// architecture.  Every option it may set leaves the computed bits unchanged
// (strip widths, workgroup sizes, wavefronts per workgroup), so a table only
// moves time.  Options set explicitly with oks_ctx_set_option win over it.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace oks {

struct TuneEntry {
    std::string op;
    std::vector<int64_t> key;
    std::map<std::string, int64_t> opts;
};

struct TuneTable {
    std::string arch;    // e.g. "gfx950"
    std::string source;  // file it was read from
    std::vector<TuneEntry> entries;
    const TuneEntry* find(const std::string& op, const std::vector<int64_t>& key) const;
};

// Parses the table's JSON text; throws std::runtime_error on malformed input.
TuneTable parse_tune_table(const std::string& text);

// The process-wide table: $OKS_TUNE_FILE if set (empty = none), otherwise
// tuning/<arch>.json beside the library, loaded once per architecture.
// Returns nullptr when there is no table for `arch`.
const TuneTable* tune_table_for(const std::string& arch);

}  // namespace oks
