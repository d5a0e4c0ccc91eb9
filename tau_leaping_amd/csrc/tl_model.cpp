// tl_model.cpp -- StochKit XML reader, channel-major assembly and the derived
// layouts of a model, plus the model part of the C ABI.
//
// Reference: StochKitReader::ReadXML (tau_leaping/StochKitReader.cpp:35-179),
// ReadXMLToMatrixStruct (tau_leaping/CSCHostMatrixData.cpp:19-107),
// OrderDetectorAndIndFiller (tau_leaping/CSCMatrix.cu:185-251),
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

// ---------------------------------------------------------------- XML ----
// A small DOM for the StochKit subset: elements, attributes, character data,
// comments, processing instructions and declarations (skipped).
namespace {

struct XNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string text;
    std::vector<std::unique_ptr<XNode>> kids;

    const XNode *child(const char *n) const {
        for (auto &k : kids)
            if (k->name == n) return k.get();
        return nullptr;
    }
    const std::string *attr(const char *n) const {
        for (auto &a : attrs)
            if (a.first == n) return &a.second;
        return nullptr;
    }
};

std::string trim(const std::string &s) {
    size_t b = 0, e = s.size();
    while (b < e && isspace((unsigned char)s[b])) ++b;
    while (e > b && isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

std::string decode_entities(const std::string &s) {
    std::string out;
    out.reserve(s.size());
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] != '&') {
            out += s[i];
            continue;
        }
        size_t semi = s.find(';', i);
        if (semi == std::string::npos) {
            out += s[i];
            continue;
        }
        std::string ent = s.substr(i + 1, semi - i - 1);
        if (ent == "lt") out += '<';
        else if (ent == "gt") out += '>';
        else if (ent == "amp") out += '&';
        else if (ent == "quot") out += '"';
        else if (ent == "apos") out += '\'';
        else if (!ent.empty() && ent[0] == '#') {
            long v = (ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X')) ? strtol(ent.c_str() + 2, nullptr, 16)
                                                                          : strtol(ent.c_str() + 1, nullptr, 10);
            if (v > 0 && v < 128) out += (char)v;
        } else {
            out += s.substr(i, semi - i + 1);
        }
        i = semi;
    }
    return out;
}

class XmlParser {
public:
    XmlParser(const std::string &t) : s_(t) {}

    bool parse(std::vector<std::unique_ptr<XNode>> &roots, std::string &err) {
        while (true) {
            skip_misc();
            if (p_ >= s_.size()) return true;
            if (s_[p_] != '<') {
                err = "unexpected character data outside an element";
                return false;
            }
            auto n = std::make_unique<XNode>();
            if (!element(*n, err)) return false;
            roots.push_back(std::move(n));
        }
    }

private:
    const std::string &s_;
    size_t p_ = 0;

    bool starts(const char *lit) const { return s_.compare(p_, strlen(lit), lit) == 0; }

    // whitespace, comments, <? ?>, <! > declarations
    void skip_misc() {
        for (;;) {
            while (p_ < s_.size() && isspace((unsigned char)s_[p_])) ++p_;
            if (starts("<!--")) {
                size_t e = s_.find("-->", p_ + 4);
                p_ = (e == std::string::npos) ? s_.size() : e + 3;
            } else if (starts("<?")) {
                size_t e = s_.find("?>", p_ + 2);
                p_ = (e == std::string::npos) ? s_.size() : e + 2;
            } else if (starts("<!")) {
                size_t e = s_.find('>', p_ + 2);
                p_ = (e == std::string::npos) ? s_.size() : e + 1;
            } else {
                return;
            }
        }
    }

    bool name_tok(std::string &out) {
        size_t b = p_;
        while (p_ < s_.size() && !isspace((unsigned char)s_[p_]) && s_[p_] != '>' && s_[p_] != '/' && s_[p_] != '=')
            ++p_;
        out = s_.substr(b, p_ - b);
        return !out.empty();
    }

    void ws() {
        while (p_ < s_.size() && isspace((unsigned char)s_[p_])) ++p_;
    }

    bool element(XNode &n, std::string &err) {
        ++p_;  // '<'
        if (!name_tok(n.name)) {
            err = "empty element name";
            return false;
        }
        for (;;) {
            ws();
            if (p_ >= s_.size()) {
                err = "unterminated tag <" + n.name;
                return false;
            }
            if (s_[p_] == '/') {
                if (p_ + 1 < s_.size() && s_[p_ + 1] == '>') {
                    p_ += 2;
                    return true;
                }
                err = "bad '/' in tag <" + n.name;
                return false;
            }
            if (s_[p_] == '>') {
                ++p_;
                break;
            }
            std::string an;
            if (!name_tok(an)) {
                err = "bad attribute in <" + n.name;
                return false;
            }
            ws();
            if (p_ >= s_.size() || s_[p_] != '=') {
                err = "attribute without value in <" + n.name;
                return false;
            }
            ++p_;
            ws();
            if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\'')) {
                err = "unquoted attribute in <" + n.name;
                return false;
            }
            char q = s_[p_++];
            size_t e = s_.find(q, p_);
            if (e == std::string::npos) {
                err = "unterminated attribute in <" + n.name;
                return false;
            }
            n.attrs.emplace_back(an, decode_entities(s_.substr(p_, e - p_)));
            p_ = e + 1;
        }
        // content
        std::string text;
        for (;;) {
            if (p_ >= s_.size()) {
                err = "missing </" + n.name + ">";
                return false;
            }
            if (starts("<!--")) {
                size_t e = s_.find("-->", p_ + 4);
                p_ = (e == std::string::npos) ? s_.size() : e + 3;
                continue;
            }
            if (starts("<![CDATA[")) {
                size_t e = s_.find("]]>", p_ + 9);
                if (e == std::string::npos) {
                    err = "unterminated CDATA";
                    return false;
                }
                text += s_.substr(p_ + 9, e - p_ - 9);
                p_ = e + 3;
                continue;
            }
            if (starts("<?")) {
                size_t e = s_.find("?>", p_ + 2);
                p_ = (e == std::string::npos) ? s_.size() : e + 2;
                continue;
            }
            if (starts("</")) {
                p_ += 2;
                std::string cn;
                name_tok(cn);
                ws();
                if (cn != n.name || p_ >= s_.size() || s_[p_] != '>') {
                    err = "mismatched </" + cn + "> for <" + n.name + ">";
                    return false;
                }
                ++p_;
                n.text = decode_entities(trim(text));
                return true;
            }
            if (s_[p_] == '<') {
                auto k = std::make_unique<XNode>();
                if (!element(*k, err)) return false;
                n.kids.push_back(std::move(k));
                continue;
            }
            text += s_[p_++];
        }
    }
};

const XNode *need(const XNode *parent, const char *name, std::string &msg) {
    const XNode *n = parent->child(name);
    if (!n) msg = std::string("can not find ") + name + " node";
    return n;
}

}  // namespace

tl_status parse_stochkit(const std::string &text, StochKitDoc &doc, std::string &msg) {
    std::vector<std::unique_ptr<XNode>> roots;
    XmlParser xp(text);
    if (!xp.parse(roots, msg)) return TL_ERR_PARSE;
    const XNode *model = nullptr;
    for (auto &r : roots)
        if (r->name == "Model") {
            model = r.get();
            break;
        }
    if (!model) {
        msg = "can not find Model node";
        return TL_ERR_PARSE;
    }
    const XNode *n;
    if (!(n = need(model, "NumberOfReactions", msg))) return TL_ERR_PARSE;
    doc.n_reactions_declared = atoi(n->text.c_str());
    if (!(n = need(model, "NumberOfSpecies", msg))) return TL_ERR_PARSE;
    doc.n_species_declared = atoi(n->text.c_str());

    // ParametersList / Parameter {Id, Expression}: numeric expressions only
    const XNode *plist = need(model, "ParametersList", msg);
    if (!plist) return TL_ERR_PARSE;
    bool any = false;
    for (auto &p : plist->kids) {
        if (p->name != "Parameter") continue;
        any = true;
        const XNode *id = need(p.get(), "Id", msg), *ex = id ? need(p.get(), "Expression", msg) : nullptr;
        if (!id || !ex) return TL_ERR_PARSE;
        char *end = nullptr;
        double v = strtod(ex->text.c_str(), &end);
        if (end == ex->text.c_str()) {
            msg = "parameter " + id->text + ": expression '" + ex->text + "' is not a number";
            return TL_ERR_PARSE;
        }
        doc.params[id->text] = v;
    }
    if (!any) {
        msg = "can not find Parameter node";
        return TL_ERR_PARSE;
    }

    // ReactionsList / Reaction
    const XNode *rlist = need(model, "ReactionsList", msg);
    if (!rlist) return TL_ERR_PARSE;
    any = false;
    for (auto &r : rlist->kids) {
        if (r->name != "Reaction") continue;
        any = true;
        ReactionDesc rd;
        const XNode *rate = need(r.get(), "Rate", msg);
        if (!rate) return TL_ERR_PARSE;
        rd.rate_id = rate->text;
        auto it = doc.params.find(rd.rate_id);
        if (it == doc.params.end()) {
            msg = "can not find the rate for the reaction " + rd.rate_id;
            return TL_ERR_PARSE;
        }
        rd.rate = it->second;
        const XNode *reac = need(r.get(), "Reactants", msg);
        if (!reac) return TL_ERR_PARSE;
        const XNode *prod = r->child("Products");
        for (int side = 0; side < 2; ++side) {
            const XNode *lst = side == 0 ? reac : prod;
            if (!lst) continue;
            for (auto &sr : lst->kids) {
                if (sr->name != "SpeciesReference") continue;
                const std::string *id = sr->attr("id"), *st = sr->attr("stoichiometry");
                if (!id || !st) {
                    msg = "SpeciesReference without id or stoichiometry";
                    return TL_ERR_PARSE;
                }
                (side == 0 ? rd.reactants : rd.products)[*id] = atoi(st->c_str());
            }
        }
        doc.reactions.push_back(std::move(rd));
    }
    if (!any) {
        msg = "can not find Reaction node";
        return TL_ERR_PARSE;
    }
    if ((int32_t)doc.reactions.size() != doc.n_reactions_declared) {
        msg = "reaction count mismatch: " + std::to_string(doc.reactions.size()) + " read, " +
              std::to_string(doc.n_reactions_declared) + " declared";
        return TL_ERR_PARSE;
    }

    // SpeciesList / Species {Id, InitialPopulation}
    const XNode *slist = need(model, "SpeciesList", msg);
    if (!slist) return TL_ERR_PARSE;
    std::unordered_map<std::string, size_t> seen;
    for (auto &s : slist->kids) {
        if (s->name != "Species") continue;
        const XNode *id = need(s.get(), "Id", msg), *pop = id ? need(s.get(), "InitialPopulation", msg) : nullptr;
        if (!id || !pop) return TL_ERR_PARSE;
        auto f = seen.find(id->text);
        if (f != seen.end()) {  // a repeated Id keeps its first index, takes the last population
            doc.populations[f->second] = atoi(pop->text.c_str());
            continue;
        }
        seen[id->text] = doc.species.size();
        doc.species.push_back(id->text);
        doc.populations.push_back(atoi(pop->text.c_str()));
    }
    if (doc.species.empty()) {
        msg = "can not find Species node";
        return TL_ERR_PARSE;
    }
    if ((int32_t)doc.species.size() != doc.n_species_declared) {
        msg = "species count mismatch: " + std::to_string(doc.species.size()) + " read, " +
              std::to_string(doc.n_species_declared) + " declared";
        return TL_ERR_PARSE;
    }
    return TL_OK;
}

// ------------------------------------------------------------- layouts ----

// Reaction order and reactant indices from the negative entries of column j in
// species order.  Accepted patterns: S-> (1), Sa+Sb-> (2), 2S-> (3).
static bool channel_order(const tl_model &m, int32_t j, int8_t &order, int32_t &a, int32_t &b) {
    int nneg = 0, v0 = 0, v1 = 0;
    int32_t i0 = -1, i1 = -1;
    for (int32_t e = m.col_ptr[j]; e < m.col_ptr[j + 1]; ++e) {
        if (m.stoich[e] >= 0) continue;
        if (nneg == 0) {
            v0 = m.stoich[e];
            i0 = m.row_ind[e];
        } else {
            v1 = m.stoich[e];
            i1 = m.row_ind[e];
        }
        ++nneg;
    }
    if (nneg == 1 && v0 == -1) {
        order = 1;
        a = b = i0;
        return true;
    }
    if (nneg == 1 && v0 == -2) {
        order = 3;
        a = b = i0;
        return true;
    }
    if (nneg == 2 && v0 == -1 && v1 == -1 && i0 != i1) {
        order = 2;
        a = i0;
        b = i1;
        return true;
    }
    return false;
}

tl_status build_model(tl_model &m, std::string &msg) {
    if (m.N <= 0 || m.M <= 0) {
        msg = "model needs at least one species and one channel";
        return TL_ERR_INVALID_ARG;
    }
    if ((int64_t)m.M * 8 > INT32_MAX || (int64_t)m.N * 8 > INT32_MAX) {
        msg = "model too large for the packed 29-bit index";
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
        if (m.stoich[e] < -2 || m.stoich[e] > 5) {
            msg = "stoichiometry " + std::to_string((int)m.stoich[e]) + " outside [-2, 5] at channel entry " +
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
    for (int32_t j = 0; j < m.M; ++j)
        if (!channel_order(m, j, m.order[j], m.s0[j], m.s1[j])) {
            msg = "channel " + std::to_string(j) + ": unsupported reactant pattern (need S->, 2S-> or Sa+Sb->)";
            return TL_ERR_UNSUPPORTED_ORDER;
        }
    // species-major transpose, rows in channel order
    m.row_ptr.assign(m.N + 1, 0);
    for (int32_t e = 0; e < m.nnz; ++e) m.row_ptr[m.row_ind[e] + 1]++;
    for (int32_t i = 0; i < m.N; ++i) m.row_ptr[i + 1] += m.row_ptr[i];
    std::vector<int32_t> fill(m.row_ptr.begin(), m.row_ptr.end() - 1);
    m.csr.assign(m.nnz, 0);
    m.csc.assign(m.nnz, 0);
    for (int32_t j = 0; j < m.M; ++j)
        for (int32_t e = m.col_ptr[j]; e < m.col_ptr[j + 1]; ++e) {
            const int32_t v3 = (m.stoich[e] + 2) & 7;
            m.csr[fill[m.row_ind[e]]++] = (j << 3) | v3;
            m.csc[e] = (m.row_ind[e] << 3) | v3;
        }
    m.g_order.assign(m.N, 0);
    for (int32_t i = 0; i < m.N && i < m.M; ++i) m.g_order[i] = m.order[i];
    // reactant-of lists: the channels whose propensity changes with x_i
    m.dep_ptr.assign(m.N + 1, 0);
    for (int32_t j = 0; j < m.M; ++j) {
        m.dep_ptr[m.s0[j] + 1]++;
        if (m.s1[j] != m.s0[j]) m.dep_ptr[m.s1[j] + 1]++;
    }
    for (int32_t i = 0; i < m.N; ++i) m.dep_ptr[i + 1] += m.dep_ptr[i];
    m.dep.assign(m.dep_ptr[m.N], 0);
    {
        std::vector<int32_t> f(m.dep_ptr.begin(), m.dep_ptr.end() - 1);
        for (int32_t j = 0; j < m.M; ++j) {
            m.dep[f[m.s0[j]]++] = j;
            if (m.s1[j] != m.s0[j]) m.dep[f[m.s1[j]]++] = j;
        }
    }
    // fixed-width columns
    int32_t max_col = 0;
    for (int32_t j = 0; j < m.M; ++j) max_col = std::max(max_col, m.col_ptr[j + 1] - m.col_ptr[j]);
    m.colq.clear();
    if (max_col <= 4) {
        m.colq.assign((size_t)m.M * 4, 2);
        for (int32_t j = 0; j < m.M; ++j)
            for (int32_t e = m.col_ptr[j]; e < m.col_ptr[j + 1]; ++e) m.colq[(size_t)j * 4 + (e - m.col_ptr[j])] = m.csc[e];
    }
    // affected-channel lists (union of the reactant-of lists of the column)
    {
        constexpr size_t kMaxAffected = (size_t)1 << 24;
        std::vector<int32_t> stamp(m.M, -1);
        std::vector<int32_t> ptr(m.M + 1, 0), list;
        bool ok = true;
        for (int32_t j = 0; j < m.M && ok; ++j) {
            const size_t start = list.size();
            for (int32_t e = m.col_ptr[j]; e < m.col_ptr[j + 1]; ++e) {
                const int32_t i = m.row_ind[e];
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
    // row offsets at propensity-tile boundaries (rows are sorted by channel)
    {
        constexpr int32_t kTile = 4096;
        const int32_t nt = std::max(1, (m.M + kTile - 1) / kTile);
        m.tile_ptr.assign((size_t)m.N * (nt + 1), 0);
        for (int32_t i = 0; i < m.N; ++i) {
            int32_t e = m.row_ptr[i];
            for (int32_t t = 0; t <= nt; ++t) {
                const int32_t bound = (t == nt) ? INT32_MAX : t * kTile;
                while (e < m.row_ptr[i + 1] && (m.csr[e] >> 3) < bound) ++e;
                m.tile_ptr[(size_t)i * (nt + 1) + t] = e;
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

// Channel-major assembly (ReadXMLToMatrixStruct): entries sorted by species;
// a species in both lists yields a reactant entry then a product entry.
static tl_status assemble(const StochKitDoc &doc, tl_model &m, std::string &msg) {
    std::unordered_map<std::string, int32_t> idx;
    for (size_t i = 0; i < doc.species.size(); ++i) idx[doc.species[i]] = (int32_t)i;
    m.N = (int32_t)doc.species.size();
    m.M = (int32_t)doc.reactions.size();
    m.col_ptr.assign(1, 0);
    m.row_ind.clear();
    m.stoich.clear();
    m.rates.clear();
    for (const ReactionDesc &r : doc.reactions) {
        std::vector<int32_t> ind;
        for (auto &kv : r.reactants) {
            auto f = idx.find(kv.first);
            if (f == idx.end()) {
                msg = "reaction uses undeclared species " + kv.first;
                return TL_ERR_PARSE;
            }
            ind.push_back(f->second);
        }
        for (auto &kv : r.products) {
            auto f = idx.find(kv.first);
            if (f == idx.end()) {
                msg = "reaction uses undeclared species " + kv.first;
                return TL_ERR_PARSE;
            }
            ind.push_back(f->second);
        }
        std::sort(ind.begin(), ind.end());
        std::set<int32_t> found;
        for (int32_t s : ind) {
            const std::string &nm = doc.species[s];
            auto rit = r.reactants.find(nm);
            int v;
            if (rit != r.reactants.end() && !found.count(s)) {
                v = -rit->second;
                found.insert(s);
            } else {
                v = r.products.at(nm);
            }
            if (v < -128 || v > 127) {
                msg = "stoichiometry out of range";
                return TL_ERR_STOICH_RANGE;
            }
            m.row_ind.push_back(s);
            m.stoich.push_back((int8_t)v);
        }
        m.col_ptr.push_back((int32_t)m.row_ind.size());
        m.rates.push_back(r.rate);
    }
    m.x0 = doc.populations;
    m.names = doc.species;
    return TL_OK;
}

}  // namespace tl

tl_model::~tl_model() { tl::release_device_copies(*this); }

using namespace tl;

extern "C" {

tl_status tl_model_from_stochkit(const char *path, tl_model **out) {
    if (!out || !path) return set_error(TL_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) return set_error(TL_ERR_IO, std::string("can not open file ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    StochKitDoc doc;
    std::string msg;
    tl_status st = parse_stochkit(ss.str(), doc, msg);
    if (st != TL_OK) return set_error(st, std::string(path) + ": " + msg);
    auto m = std::make_unique<tl_model>();
    if ((st = assemble(doc, *m, msg)) != TL_OK) return set_error(st, std::string(path) + ": " + msg);
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
