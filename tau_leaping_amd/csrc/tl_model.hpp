// tl_model.hpp -- host-side model: CSC entries as read, reaction orders and
// the packed species-major / channel-major layouts derived from them.
#pragma once

#include <stdint.h>
#include <stdio.h>

#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "tl_api.h"

struct tl_model;

namespace tl {

struct DeviceModelCache;  // device copies, one per HIP device (tl_engine)

// Read a StochKit 2.0 file into m: N, M, CSC entries, rates, x0, names
// (TL_OK, TL_ERR_PARSE or TL_ERR_STOICH_RANGE with msg set).
tl_status read_stochkit(FILE *f, tl_model &m, std::string &msg);
}  // namespace tl

struct tl_model {
    int32_t N = 0, M = 0, nnz = 0;
    std::vector<int32_t> col_ptr, row_ind;
    std::vector<int8_t> stoich;
    std::vector<double> rates;
    std::vector<int32_t> x0;
    std::vector<std::string> names;
    std::unordered_map<std::string, int32_t> name_index;

    // Net stoichiometry in the layouts below: the entries of one species in
    // one channel (a catalyst: a reactant and a product entry) merge into one
    // entry holding their sum, dropped when the sum is 0.  Off: the entries
    // as read, so a catalyst adds v^2 a_j twice to sigma^2.  Reaction orders
    // come from the entries as read either way.
    bool net_stoich = false;

    // derived
    std::vector<int32_t> lcol_ptr;  // [M+1] column offsets of csc (col_ptr unless net_stoich)
    int32_t lnnz = 0;               // entries of csr / csc
    std::vector<int8_t> order;
    std::vector<int32_t> s0, s1;
    std::vector<int32_t> s2;  // third reactant of order-4 channels; empty unless an order 0, 4, 5 or 6 exists
    std::vector<int32_t> row_ptr, csr;  // species-major, (channel << 4) | (v + 3)
    std::vector<int32_t> csc;           // channel-major, (species << 4) | (v + 3)
    std::vector<int8_t> g_order;        // order[i] for i < M else 0
    std::vector<int32_t> dep_ptr, dep;  // species i -> channels with i as a reactant
    // [M][4] column entries padded with (0 << 4) | 3 (v = 0); empty unless
    // every column has at most 4 entries
    std::vector<int32_t> colq;
    // channel j -> channels with a reactant among j's column species (the
    // propensities one firing of j can change); empty when it would exceed
    // kMaxAffected entries
    std::vector<int32_t> aff_ptr, aff;
    // [n_grains + 1][N]: first CSR entry of row i whose channel is >= g * 4096
    // (n_grains = ceil(M / 4096); grain n_grains holds the row ends)
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
