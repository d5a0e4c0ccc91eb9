// tl_model.hpp -- host-side model: StochKit reader, channel-major assembly,
// reaction orders and the packed species-major / channel-major layouts.
#pragma once

#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "tl_api.h"

namespace tl {

struct DeviceModelCache;  // device copies, one per HIP device (tl_engine)

// One reaction as read from StochKit XML (StochKitReader.cpp:35-90):
// species name -> stoichiometry for reactants and products.
struct ReactionDesc {
    std::string rate_id;
    double rate = 0.0;
    std::map<std::string, int> reactants;
    std::map<std::string, int> products;
};

struct StochKitDoc {
    int32_t n_reactions_declared = 0;
    int32_t n_species_declared = 0;
    std::unordered_map<std::string, double> params;
    std::vector<ReactionDesc> reactions;
    std::vector<std::string> species;       // SpeciesList order = species index
    std::vector<int32_t> populations;
};

// Parse StochKit 2.0 XML text.  Returns TL_OK or TL_ERR_PARSE with msg set.
tl_status parse_stochkit(const std::string &text, StochKitDoc &doc, std::string &msg);

}  // namespace tl

struct tl_model {
    int32_t N = 0, M = 0, nnz = 0;
    std::vector<int32_t> col_ptr, row_ind;
    std::vector<int8_t> stoich;
    std::vector<double> rates;
    std::vector<int32_t> x0;
    std::vector<std::string> names;
    std::unordered_map<std::string, int32_t> name_index;

    // derived
    std::vector<int8_t> order;
    std::vector<int32_t> s0, s1;
    std::vector<int32_t> row_ptr, csr;  // species-major, (channel << 3) | ((v + 2) & 7)
    std::vector<int32_t> csc;           // channel-major, (species << 3) | ((v + 2) & 7)
    std::vector<int8_t> g_order;        // order[i] for i < M else 0
    std::vector<int32_t> dep_ptr, dep;  // species i -> channels with i as a reactant
    // [M][4] column entries padded with (0 << 3) | 2 (v = 0); empty unless
    // every column has at most 4 entries
    std::vector<int32_t> colq;
    // channel j -> channels with a reactant among j's column species (the
    // propensities one firing of j can change); empty when it would exceed
    // kMaxAffected entries
    std::vector<int32_t> aff_ptr, aff;
    // [N][n_grains + 1]: first CSR entry of row i whose channel is >= g * 4096
    // (n_grains = ceil(M / 4096); entry n_grains is the row end)
    std::vector<int32_t> tile_ptr;

    uint64_t version = 1;  // bumped by set_rates / set_x0; invalidates device copies
    mutable std::mutex dev_mu;
    mutable std::map<int, tl::DeviceModelCache *> dev;

    ~tl_model();
};

namespace tl {
// Validate the channel-major entries and build the derived layouts.
tl_status build_model(tl_model &m, std::string &msg);
}  // namespace tl
