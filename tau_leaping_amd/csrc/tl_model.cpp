// tl_model.cpp -- the derived layouts of a model (reaction orders, packed
// species-major and channel-major maps, dependency and affected-channel
// lists, tiles) and the model part of the C ABI.  The StochKit reader is in
// tl_stochkit.cpp.
//
// Reference: OrderDetectorAndIndFiller (tau_leaping/CSCMatrix.cu:185-251),
// ToCSR / CSRMatrixPacked ctor (CSCMatrix.cu:147-178, CSRMatrixPacked.cu:22-31).
#include "tl_model.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <set>
#include <sstream>

#include "tl_internal.hpp"

namespace tl {

// ------------------------------------------------------------- layouts ----

// Reaction order and reactant indices from the negative entries of column j in
// species order.  Accepted patterns: -> (0, a source), S-> (1), Sa+Sb-> (2),
// 2S-> (3), and three reactant molecules: Sa+Sb+Sc-> (4, third index in c),
// 2Sa+Sb-> (5, a the species taken twice), 3S-> (6).
static bool channel_order(const tl_model &m, int32_t j, int8_t &order, int32_t &a, int32_t &b, int32_t &c) {
    int nneg = 0;
    int v[3] = {0, 0, 0};
    int32_t id[3] = {-1, -1, -1};
    for (int32_t e = m.col_ptr[j]; e < m.col_ptr[j + 1]; ++e) {
        if (m.stoich[e] >= 0) continue;
        if (nneg < 3) {
            v[nneg] = m.stoich[e];
            id[nneg] = m.row_ind[e];
        }
        ++nneg;
    }
    a = b = c = 0;
    if (nneg == 0) {  // a source: no reactant, a = c
        order = 0;
        return true;
    }
    if (nneg == 1 && v[0] == -1) {
        order = 1;
        a = b = id[0];
        return true;
    }
    if (nneg == 1 && v[0] == -2) {
        order = 3;
        a = b = id[0];
        return true;
    }
    if (nneg == 1 && v[0] == -3) {
        order = 6;  // 3S
        a = b = id[0];
        return true;
    }
    if (nneg == 2 && v[0] == -1 && v[1] == -1 && id[0] != id[1]) {
        order = 2;
        a = id[0];
        b = id[1];
        return true;
    }
    if (nneg == 2 && id[0] != id[1] && (v[0] + v[1] == -3) && (v[0] == -2 || v[1] == -2)) {
        order = 5;  // 2Sa + Sb: a is the species taken twice
        a = v[0] == -2 ? id[0] : id[1];
        b = v[0] == -2 ? id[1] : id[0];
        return true;
    }
    if (nneg == 3 && v[0] == -1 && v[1] == -1 && v[2] == -1 && id[0] != id[1] && id[1] != id[2] && id[0] != id[2]) {
        order = 4;  // Sa + Sb + Sc
        a = id[0];
        b = id[1];
        c = id[2];
        return true;
    }
    return false;
}

tl_status build_model(tl_model &m, std::string &msg) {
    if (m.N <= 0 || m.M <= 0) {
        msg = "model needs at least one species and one channel";
        return TL_ERR_INVALID_ARG;
    }
    if ((int64_t)m.M * 16 > INT32_MAX || (int64_t)m.N * 16 > INT32_MAX) {
        msg = "model too large for the packed 27-bit index";
        return TL_ERR_TOO_LARGE;
    }
    if ((int32_t)m.col_ptr.size() != m.M + 1 || m.col_ptr[0] != 0) {
        msg = "col_ptr must have M+1 entries starting at 0";
        return TL_ERR_INVALID_ARG;
    }
    for (int32_t j = 0; j < m.M; ++j)
        if (m.col_ptr[j + 1] < m.col_ptr[j]) {
            msg = "col_ptr is not monotone";
            return TL_ERR_INVALID_ARG;
        }
    m.nnz = m.col_ptr[m.M];
    if ((int32_t)m.row_ind.size() < m.nnz || (int32_t)m.stoich.size() < m.nnz) {
        msg = "row_ind / stoich shorter than col_ptr[M]";
        return TL_ERR_INVALID_ARG;
    }
    for (int32_t e = 0; e < m.nnz; ++e) {
        if (m.row_ind[e] < 0 || m.row_ind[e] >= m.N) {
            msg = "species index out of range at entry " + std::to_string(e);
            return TL_ERR_INVALID_ARG;
        }
        if (m.stoich[e] < -3 || m.stoich[e] > 12) {
            msg = "stoichiometry " + std::to_string((int)m.stoich[e]) + " outside [-3, 12] at channel entry " +
                  std::to_string(e);
            return TL_ERR_STOICH_RANGE;
        }
    }
    for (int32_t j = 0; j < m.M; ++j)
        if (!(m.rates[j] >= 0.0)) {
            msg = "rate of channel " + std::to_string(j) + " is negative or NaN";
            return TL_ERR_INVALID_ARG;
        }
    for (int32_t i = 0; i < m.N; ++i)
        if (m.x0[i] < 0) {
            msg = "initial population of species " + std::to_string(i) + " is negative";
            return TL_ERR_INVALID_ARG;
        }
    m.order.assign(m.M, 0);
    m.s0.assign(m.M, 0);
    m.s1.assign(m.M, 0);
    std::vector<int32_t> third(m.M, 0);
    bool ext = false;
    for (int32_t j = 0; j < m.M; ++j) {
        if (!channel_order(m, j, m.order[j], m.s0[j], m.s1[j], third[j])) {
            msg = "channel " + std::to_string(j) +
                  ": unsupported reactant pattern (need ->, S->, 2S->, Sa+Sb->, Sa+Sb+Sc->, 2Sa+Sb-> or 3S->)";
            return TL_ERR_UNSUPPORTED_ORDER;
        }
        ext |= m.order[j] == 0 || m.order[j] >= 4;
    }
    // third reactant of Sa+Sb+Sc channels; empty unless some channel is a
    // source or has three reactant molecules (the kernels then take the
    // extended formulas)
    if (ext) m.s2.swap(third);
    else m.s2.clear();
    // layout entries: as read, or one entry per (channel, species) holding
    // the net stoichiometry, in species order, zero sums dropped
    std::vector<int32_t> net_row;
    std::vector<int8_t> net_v;
    const int32_t *lrow = m.row_ind.data();
    const int8_t *lv = m.stoich.data();
    if (m.net_stoich) {
        m.lcol_ptr.assign(m.M + 1, 0);
        net_row.reserve(m.nnz);
        net_v.reserve(m.nnz);
        std::vector<std::pair<int32_t, int32_t>> col;
        for (int32_t j = 0; j < m.M; ++j) {
            col.clear();
            for (int32_t e = m.col_ptr[j]; e < m.col_ptr[j + 1]; ++e) col.emplace_back(m.row_ind[e], m.stoich[e]);
            std::stable_sort(col.begin(), col.end(),
                             [](const std::pair<int32_t, int32_t> &x, const std::pair<int32_t, int32_t> &y) {
                                 return x.first < y.first;
                             });
            for (size_t k = 0; k < col.size();) {
                int32_t v = 0;
                size_t k1 = k;
                for (; k1 < col.size() && col[k1].first == col[k].first; ++k1) v += col[k1].second;
                if (v < -3 || v > 12) {
                    msg = "net stoichiometry " + std::to_string(v) + " of species " + std::to_string(col[k].first) +
                          " in channel " + std::to_string(j) + " outside [-3, 12]";
                    return TL_ERR_STOICH_RANGE;
                }
                if (v != 0) {
                    net_row.push_back(col[k].first);
                    net_v.push_back((int8_t)v);
                }
                k = k1;
            }
            m.lcol_ptr[j + 1] = (int32_t)net_row.size();
        }
        lrow = net_row.data();
        lv = net_v.data();
    } else {
        m.lcol_ptr.assign(m.col_ptr.begin(), m.col_ptr.begin() + m.M + 1);
    }
    m.lnnz = m.lcol_ptr[m.M];
    const std::vector<int32_t> &lcol = m.lcol_ptr;
    // species-major transpose, rows in channel order
    m.row_ptr.assign(m.N + 1, 0);
    for (int32_t e = 0; e < m.lnnz; ++e) m.row_ptr[lrow[e] + 1]++;
    for (int32_t i = 0; i < m.N; ++i) m.row_ptr[i + 1] += m.row_ptr[i];
    std::vector<int32_t> fill(m.row_ptr.begin(), m.row_ptr.end() - 1);
    m.csr.assign(m.lnnz, 0);
    m.csc.assign(m.lnnz, 0);
    for (int32_t j = 0; j < m.M; ++j)
        for (int32_t e = lcol[j]; e < lcol[j + 1]; ++e) {
            const int32_t v4 = lv[e] + 3;
            m.csr[fill[lrow[e]]++] = (j << 4) | v4;
            m.csc[e] = (lrow[e] << 4) | v4;
        }
    m.g_order.assign(m.N, 0);
    for (int32_t i = 0; i < m.N && i < m.M; ++i) m.g_order[i] = m.order[i];
    // reactant-of lists: the channels whose propensity changes with x_i
    m.dep_ptr.assign(m.N + 1, 0);
    for (int32_t j = 0; j < m.M; ++j) {
        if (m.order[j] == 0) continue;  // sources read no species
        m.dep_ptr[m.s0[j] + 1]++;
        if (m.s1[j] != m.s0[j]) m.dep_ptr[m.s1[j] + 1]++;
        if (m.order[j] == 4) m.dep_ptr[m.s2[j] + 1]++;
    }
    for (int32_t i = 0; i < m.N; ++i) m.dep_ptr[i + 1] += m.dep_ptr[i];
    m.dep.assign(m.dep_ptr[m.N], 0);
    {
        std::vector<int32_t> f(m.dep_ptr.begin(), m.dep_ptr.end() - 1);
        for (int32_t j = 0; j < m.M; ++j) {
            if (m.order[j] == 0) continue;
            m.dep[f[m.s0[j]]++] = j;
            if (m.s1[j] != m.s0[j]) m.dep[f[m.s1[j]]++] = j;
            if (m.order[j] == 4) m.dep[f[m.s2[j]]++] = j;
        }
    }
    // fixed-width columns
    int32_t max_col = 0;
    for (int32_t j = 0; j < m.M; ++j) max_col = std::max(max_col, lcol[j + 1] - lcol[j]);
    m.colq.clear();
    if (max_col <= 4) {
        m.colq.assign((size_t)m.M * 4, 3);  // padding: species 0, v = 0
        for (int32_t j = 0; j < m.M; ++j)
            for (int32_t e = lcol[j]; e < lcol[j + 1]; ++e) m.colq[(size_t)j * 4 + (e - lcol[j])] = m.csc[e];
    }
    // affected-channel lists (union of the reactant-of lists of the column)
    {
        constexpr size_t kMaxAffected = (size_t)1 << 24;
        std::vector<int32_t> stamp(m.M, -1);
        std::vector<int32_t> ptr(m.M + 1, 0), list;
        bool ok = true;
        for (int32_t j = 0; j < m.M && ok; ++j) {
            const size_t start = list.size();
            for (int32_t e = lcol[j]; e < lcol[j + 1]; ++e) {
                const int32_t i = lrow[e];
                for (int32_t d = m.dep_ptr[i]; d < m.dep_ptr[i + 1]; ++d)
                    if (stamp[m.dep[d]] != j) {
                        stamp[m.dep[d]] = j;
                        list.push_back(m.dep[d]);
                    }
            }
            std::sort(list.begin() + start, list.end());
            ptr[j + 1] = (int32_t)list.size();
            if (list.size() > kMaxAffected) ok = false;
        }
        m.aff_ptr.clear();
        m.aff.clear();
        if (ok) {
            m.aff_ptr.swap(ptr);
            m.aff.swap(list);
        }
    }
    // row offsets at propensity-tile boundaries (rows are sorted by channel),
    // grain-major so that a tile pass reads them coalesced over species
    {
        constexpr int32_t kTile = 4096;
        const int32_t nt = std::max(1, (m.M + kTile - 1) / kTile);
        m.tile_ptr.assign((size_t)m.N * (nt + 1), 0);
        for (int32_t i = 0; i < m.N; ++i) {
            int32_t e = m.row_ptr[i];
            for (int32_t t = 0; t <= nt; ++t) {
                const int32_t bound = (t == nt) ? INT32_MAX : t * kTile;
                while (e < m.row_ptr[i + 1] && (m.csr[e] >> 4) < bound) ++e;
                m.tile_ptr[(size_t)t * m.N + i] = e;
            }
        }
    }
    if (m.names.empty()) {
        m.names.resize(m.N);
        for (int32_t i = 0; i < m.N; ++i) m.names[i] = "S" + std::to_string(i + 1);
    }
    m.name_index.clear();
    for (int32_t i = 0; i < m.N; ++i) m.name_index.emplace(m.names[i], i);
    return TL_OK;
}

}  // namespace tl

tl_model::~tl_model() { tl::release_device_copies(*this); }

using namespace tl;

extern "C" {

tl_status tl_model_from_stochkit(const char *path, tl_model **out) {
    if (!out || !path) return set_error(TL_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return set_error(TL_ERR_IO, std::string("can not open file ") + path);
    auto m = std::make_unique<tl_model>();
    std::string msg;
    tl_status st = read_stochkit(f, *m, msg);
    std::fclose(f);
    if (st != TL_OK) return set_error(st, std::string(path) + ": " + msg);
    if ((st = build_model(*m, msg)) != TL_OK) return set_error(st, std::string(path) + ": " + msg);
    *out = m.release();
    return TL_OK;
}

tl_status tl_model_from_csc(int32_t n_species, int32_t n_channels, const int32_t *col_ptr, const int32_t *row_ind,
                            const int8_t *stoich, const double *rates, const int32_t *x0, tl_model **out) {
    if (!out) return set_error(TL_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    if (n_species <= 0 || n_channels <= 0 || !col_ptr || !rates || !x0)
        return set_error(TL_ERR_INVALID_ARG, "invalid model arguments");
    if (col_ptr[n_channels] > 0 && (!row_ind || !stoich)) return set_error(TL_ERR_INVALID_ARG, "null entries");
    auto m = std::make_unique<tl_model>();
    m->N = n_species;
    m->M = n_channels;
    m->col_ptr.assign(col_ptr, col_ptr + n_channels + 1);
    int32_t nnz = col_ptr[n_channels];
    if (nnz < 0) return set_error(TL_ERR_INVALID_ARG, "negative nnz");
    m->row_ind.assign(row_ind, row_ind + nnz);
    m->stoich.assign(stoich, stoich + nnz);
    m->rates.assign(rates, rates + n_channels);
    m->x0.assign(x0, x0 + n_species);
    std::string msg;
    tl_status st = build_model(*m, msg);
    if (st != TL_OK) return set_error(st, msg);
    *out = m.release();
    return TL_OK;
}

void tl_model_destroy(tl_model *m) { delete m; }
int32_t tl_model_n_species(const tl_model *m) { return m ? m->N : 0; }
int32_t tl_model_n_channels(const tl_model *m) { return m ? m->M : 0; }
int32_t tl_model_nnz(const tl_model *m) { return m ? m->nnz : 0; }

const char *tl_model_species_name(const tl_model *m, int32_t i) {
    if (!m || i < 0 || i >= m->N) return nullptr;
    return m->names[i].c_str();
}

tl_status tl_model_species_index(const tl_model *m, const char *name, int32_t *out) {
    if (!m || !name || !out) return set_error(TL_ERR_INVALID_ARG, "null argument");
    auto f = m->name_index.find(name);
    if (f == m->name_index.end()) return set_error(TL_ERR_INVALID_ARG, std::string("unknown species ") + name);
    *out = f->second;
    return TL_OK;
}

tl_status tl_model_get_rates(const tl_model *m, double *rates) {
    if (!m || !rates) return set_error(TL_ERR_INVALID_ARG, "null argument");
    std::copy(m->rates.begin(), m->rates.end(), rates);
    return TL_OK;
}

tl_status tl_model_set_rates(tl_model *m, const double *rates) {
    if (!m || !rates) return set_error(TL_ERR_INVALID_ARG, "null argument");
    for (int32_t j = 0; j < m->M; ++j)
        if (!(rates[j] >= 0.0)) return set_error(TL_ERR_INVALID_ARG, "negative or NaN rate");
    std::lock_guard<std::mutex> g(m->dev_mu);
    m->rates.assign(rates, rates + m->M);
    m->version++;
    return TL_OK;
}

tl_status tl_model_get_x0(const tl_model *m, int32_t *x0) {
    if (!m || !x0) return set_error(TL_ERR_INVALID_ARG, "null argument");
    std::copy(m->x0.begin(), m->x0.end(), x0);
    return TL_OK;
}

tl_status tl_model_set_x0(tl_model *m, const int32_t *x0) {
    if (!m || !x0) return set_error(TL_ERR_INVALID_ARG, "null argument");
    for (int32_t i = 0; i < m->N; ++i)
        if (x0[i] < 0) return set_error(TL_ERR_INVALID_ARG, "negative population");
    std::lock_guard<std::mutex> g(m->dev_mu);
    m->x0.assign(x0, x0 + m->N);
    m->version++;
    return TL_OK;
}

tl_status tl_model_set_net_stoichiometry(tl_model *m, int32_t on) {
    if (!m) return set_error(TL_ERR_INVALID_ARG, "null model");
    std::lock_guard<std::mutex> g(m->dev_mu);
    const bool prev = m->net_stoich;
    if (prev == (on != 0)) return TL_OK;
    m->net_stoich = on != 0;
    std::string msg;
    const tl_status st = build_model(*m, msg);
    if (st != TL_OK) {
        m->net_stoich = prev;
        std::string again;
        build_model(*m, again);
        return set_error(st, msg);
    }
    m->version++;
    return TL_OK;
}

int32_t tl_model_net_stoichiometry(const tl_model *m) { return m && m->net_stoich ? 1 : 0; }

int32_t tl_model_layout_nnz(const tl_model *m) { return m ? m->lnnz : 0; }

tl_status tl_model_structure(const tl_model *m, int8_t *order, int32_t *s0, int32_t *s1, int32_t *row_ptr,
                             int32_t *csr) {
    if (!m) return set_error(TL_ERR_INVALID_ARG, "null model");
    if (order) std::copy(m->order.begin(), m->order.end(), order);
    if (s0) std::copy(m->s0.begin(), m->s0.end(), s0);
    if (s1) std::copy(m->s1.begin(), m->s1.end(), s1);
    if (row_ptr) std::copy(m->row_ptr.begin(), m->row_ptr.end(), row_ptr);
    if (csr) std::copy(m->csr.begin(), m->csr.end(), csr);
    return TL_OK;
}

}  // extern "C"
