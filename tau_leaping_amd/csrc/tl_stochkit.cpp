// tl_stochkit.cpp -- streaming StochKit 2.0 XML reader and channel-major
// assembly of the model it describes.
//
// Reference: StochKitReader::ReadXML (tau_leaping/StochKitReader.cpp:35-179)
// reads the document into a DOM, then walks Model / ParametersList /
// ReactionsList / SpeciesList; ReadXMLToMatrixStruct
// (tau_leaping/CSCHostMatrixData.cpp:19-107) turns the reactions into CSC
// columns sorted by species, a species that is both reactant and product
// giving a reactant entry then a product entry.
//
// This reader never holds the document: a pull tokenizer runs over a sliding
// window of the file and a small state machine keeps only what the model
// needs (interned species and rate names, one flat list of species
// references), so a 1e6-channel file (about 500 MB of text) is read in
// seconds with memory proportional to the model, not the text.
//
// Element selection is that of the DOM walk: the first root named Model; the
// first NumberOfReactions, NumberOfSpecies, ParametersList, ReactionsList and
// SpeciesList under it; every Parameter / Reaction / Species child of those
// lists; the first Id / Expression of a Parameter, the first Rate / Reactants
// / Products of a Reaction, the first Id / InitialPopulation of a Species;
// SpeciesReference children of the selected Reactants / Products.  Element
// text is the element's own character data (CDATA included), trimmed, with
// the five predefined entities and numeric references below 128 decoded.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "tl_internal.hpp"
#include "tl_model.hpp"

namespace tl {
namespace {

// ------------------------------------------------------------ tokenizer --
// Pull tokenizer over a FILE* with a sliding window.  Events: start tag (name
// and attributes; a self-closing tag also yields its end), end tag, character
// data, end of input.  Comments, processing instructions and declarations
// are skipped.
class XmlPull {
public:
    enum Ev { kStart, kEnd, kText, kEof, kError };

    explicit XmlPull(FILE *f) : f_(f), buf_(1 << 20) {}

    Ev next() {
        if (pending_end_) {
            pending_end_ = false;
            return kEnd;
        }
        for (;;) {
            const int c = peek(0);
            if (c < 0) return io_error_ ? fail("read error") : kEof;
            if (c != '<') return text_run();
            const int c1 = peek(1);
            if (c1 == '/') return end_tag();
            if (c1 == '?') {
                if (!skip_past("?>", 2)) return fail("unterminated processing instruction");
                continue;
            }
            if (c1 == '!') {
                if (starts("<!--")) {
                    if (!skip_past("-->", 4)) return fail("unterminated comment");
                    continue;
                }
                if (starts("<![CDATA[")) {
                    take(9);
                    text_.clear();
                    if (!copy_until("]]>")) return fail("unterminated CDATA");
                    tp_ = text_.data();
                    tn_ = text_.size();
                    return kText;
                }
                if (!skip_past(">", 2)) return fail("unterminated declaration");
                continue;
            }
            return start_tag();
        }
    }

    const std::string &name() const { return name_; }
    // character data of the last text event: valid until the next call
    const char *text_data() const { return tp_; }
    size_t text_size() const { return tn_; }
    const std::string &error() const { return err_; }
    const std::string *attr(const char *n) const {
        for (size_t i = 0; i < n_attrs_; ++i)
            if (attrs_[i].first == n) return &attrs_[i].second;
        return nullptr;
    }

private:
    FILE *f_;
    std::vector<char> buf_;
    size_t b_ = 0, e_ = 0;  // window [b_, e_) of buf_
    bool eof_ = false, io_error_ = false, pending_end_ = false;
    std::string name_, text_, err_;
    const char *tp_ = nullptr;
    size_t tn_ = 0;
    std::vector<std::pair<std::string, std::string>> attrs_;
    size_t n_attrs_ = 0;

    Ev fail(const std::string &m) {
        err_ = m;
        return kError;
    }

    // make at least n bytes available from b_ unless the input ends first
    bool fill(size_t n) {
        while (e_ - b_ < n && !eof_) {
            if (b_ > 0) {
                std::memmove(buf_.data(), buf_.data() + b_, e_ - b_);
                e_ -= b_;
                b_ = 0;
            }
            if (n > buf_.size() - e_) buf_.resize(std::max(buf_.size() * 2, e_ + n));
            const size_t got = std::fread(buf_.data() + e_, 1, buf_.size() - e_, f_);
            e_ += got;
            if (got == 0) {
                eof_ = true;
                io_error_ = std::ferror(f_) != 0;
            }
        }
        return e_ - b_ >= n;
    }
    int peek(size_t k) {
        if (b_ + k < e_) return (unsigned char)buf_[b_ + k];
        return fill(k + 1) ? (unsigned char)buf_[b_ + k] : -1;
    }
    void take(size_t n) { b_ += n; }
    template <size_t L>
    bool starts(const char (&lit)[L]) {
        constexpr size_t n = L - 1;
        return (e_ - b_ >= n || fill(n)) && std::memcmp(buf_.data() + b_, lit, n) == 0;
    }
    static bool space(int c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }
    void ws() {
        while (space(peek(0))) take(1);
    }

    // advance past the first occurrence of lit at or after offset `from`
    bool skip_past(const char *lit, size_t from) {
        const size_t n = std::strlen(lit);
        take(from);
        for (;;) {
            if (!fill(n)) return false;
            const char *p = buf_.data() + b_;
            const size_t avail = e_ - b_;
            const char *hit = static_cast<const char *>(memmem(p, avail, lit, n));
            if (hit) {
                take((size_t)(hit - p) + n);
                return true;
            }
            take(avail - (n - 1));  // keep a possible partial match
            if (eof_) return false;
            fill(n + 4096);
        }
    }
    // append raw bytes up to lit to text_, then consume lit
    bool copy_until(const char *lit) {
        const size_t n = std::strlen(lit);
        for (;;) {
            if (!fill(n)) return false;
            const char *p = buf_.data() + b_;
            const size_t avail = e_ - b_;
            const char *hit = static_cast<const char *>(memmem(p, avail, lit, n));
            if (hit) {
                text_.append(p, (size_t)(hit - p));
                take((size_t)(hit - p) + n);
                return true;
            }
            text_.append(p, avail - (n - 1));
            take(avail - (n - 1));
            if (eof_) return false;
            fill(n + 4096);
        }
    }

    // a run of character data up to the next '<' (or the end), returned as a
    // view of the window; the window grows if a run is longer than it
    Ev text_run() {
        size_t from = 0;
        for (;;) {
            const char *p = buf_.data() + b_;
            const size_t avail = e_ - b_;
            const char *lt = static_cast<const char *>(std::memchr(p + from, '<', avail - from));
            if (lt || !fill(avail + 1)) {
                const size_t n = lt ? (size_t)(lt - p) : e_ - b_;
                tp_ = buf_.data() + b_;
                tn_ = n;
                take(n);
                return kText;
            }
            from = avail;
        }
    }

    static bool name_end(int c) { return c < 0 || space(c) || c == '>' || c == '/' || c == '='; }
    bool name_tok(std::string &out) {
        size_t k = 0;
        while (!name_end(peek(k))) ++k;
        out.assign(buf_.data() + b_, k);
        take(k);
        return k > 0;
    }

    Ev end_tag() {
        take(2);
        name_tok(name_);
        ws();
        if (peek(0) != '>') return fail("bad end tag </" + name_);
        take(1);
        return kEnd;
    }

    Ev start_tag() {
        take(1);
        if (!name_tok(name_)) return fail("empty element name");
        n_attrs_ = 0;
        for (;;) {
            ws();
            const int c = peek(0);
            if (c < 0) return fail("unterminated tag <" + name_);
            if (c == '/') {
                if (peek(1) != '>') return fail("bad '/' in tag <" + name_);
                take(2);
                pending_end_ = true;
                return kStart;
            }
            if (c == '>') {
                take(1);
                return kStart;
            }
            if (n_attrs_ == attrs_.size()) attrs_.emplace_back();
            auto &a = attrs_[n_attrs_];
            if (!name_tok(a.first)) return fail("bad attribute in <" + name_);
            ws();
            if (peek(0) != '=') return fail("attribute without value in <" + name_);
            take(1);
            ws();
            const int q = peek(0);
            if (q != '"' && q != '\'') return fail("unquoted attribute in <" + name_);
            take(1);
            size_t k = 0;
            for (;; ++k) {
                const int v = peek(k);
                if (v < 0) return fail("unterminated attribute in <" + name_);
                if (v == q) break;
            }
            a.second.assign(buf_.data() + b_, k);
            take(k + 1);
            if (a.second.find('&') != std::string::npos) a.second = decode_entities(a.second);
            ++n_attrs_;
        }
    }

public:
    static std::string decode_entities(const std::string &s) {
        if (s.find('&') == std::string::npos) return s;
        std::string out;
        out.reserve(s.size());
        for (size_t i = 0; i < s.size(); ++i) {
            if (s[i] != '&') {
                out += s[i];
                continue;
            }
            const size_t semi = s.find(';', i);
            if (semi == std::string::npos) {
                out += s[i];
                continue;
            }
            const std::string ent = s.substr(i + 1, semi - i - 1);
            if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "amp") out += '&';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else if (!ent.empty() && ent[0] == '#') {
                const long v = (ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X'))
                                   ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                   : std::strtol(ent.c_str() + 1, nullptr, 10);
                if (v > 0 && v < 128) out += (char)v;
            } else {
                out += s.substr(i, semi - i + 1);
            }
            i = semi;
        }
        return out;
    }
};

void trim_in_place(std::string &s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    s.erase(e);
    s.erase(0, b);
}

// ---------------------------------------------------------- the document --
// What the walk keeps: rates and species are referred to by interned name.
// Name -> dense index.  Open addressing with linear probing; a slot holds
// the name's length and either the name itself (up to 8 bytes, the usual
// "S123" / "c4567") or its hash, so most lookups touch a single slot, which
// keeps a 1e6-name table fast.  Names live in one arena.
class Interner {
public:
    Interner() : slots_(1024) {}

    int32_t operator()(const char *p, size_t n) {
        const uint64_t t = tag(p, n);
        size_t i = (size_t)mix(t) & (slots_.size() - 1);
        for (;; i = (i + 1) & (slots_.size() - 1)) {
            const Slot &sl = slots_[i];
            if (sl.index < 0) break;
            if (sl.tag == t && sl.len == n && (n <= 8 || std::memcmp(arena_.data() + off_[sl.index], p, n) == 0))
                return sl.index;
        }
        const int32_t k = (int32_t)off_.size();
        off_.push_back(arena_.size());
        arena_.insert(arena_.end(), p, p + n);
        slots_[i] = Slot{t, k, (uint32_t)n};
        if (2 * off_.size() > slots_.size()) grow();
        return k;
    }
    int32_t operator()(const std::string &s) { return (*this)(s.data(), s.size()); }
    // whether name k is s (a cheap check before a lookup)
    bool is(int32_t k, const std::string &s) const {
        return k >= 0 && (size_t)k < off_.size() && len(k) == s.size() &&
               std::memcmp(arena_.data() + off_[k], s.data(), s.size()) == 0;
    }
    size_t size() const { return off_.size(); }
    std::string name(int32_t k) const { return std::string(arena_.data() + off_[k], len(k)); }

private:
    struct Slot {
        uint64_t tag = 0;
        int32_t index = -1;
        uint32_t len = 0;
    };
    std::vector<Slot> slots_;
    std::vector<char> arena_;
    std::vector<size_t> off_;

    size_t len(int32_t k) const { return ((size_t)k + 1 < off_.size() ? off_[k + 1] : arena_.size()) - off_[k]; }
    static uint64_t tag(const char *p, size_t n) {
        uint64_t h = 0;
        if (n <= 8) {
            std::memcpy(&h, p, n);
            return h;
        }
        h = 1469598103934665603ull;  // FNV-1a
        for (size_t i = 0; i < n; ++i) h = (h ^ (unsigned char)p[i]) * 1099511628211ull;
        return h;
    }
    static uint64_t mix(uint64_t h) {  // splitmix64 finaliser
        h ^= h >> 30;
        h *= 0xbf58476d1ce4e5b9ull;
        h ^= h >> 27;
        h *= 0x94d049bb133111ebull;
        return h ^ (h >> 31);
    }
    void grow() {
        std::vector<Slot> old(slots_.size() * 2);
        old.swap(slots_);
        for (const Slot &sl : old) {
            if (sl.index < 0) continue;
            size_t i = (size_t)mix(sl.tag) & (slots_.size() - 1);
            while (slots_[i].index >= 0) i = (i + 1) & (slots_.size() - 1);
            slots_[i] = sl;
        }
    }
};

struct SpeciesRef {
    int32_t name;  // interned species name
    int32_t v;     // stoichiometry as written
    int32_t side;  // 0 reactant, 1 product
};

struct Doc {
    Interner species_names, rate_names;
    bool has_nr = false, has_ns = false, has_plist = false, has_rlist = false, has_slist = false;
    int32_t n_reactions_declared = 0, n_species_declared = 0;
    bool any_param = false, any_reaction = false;
    std::vector<double> param_value;  // by interned rate name
    std::vector<char> param_set;
    std::vector<int32_t> reaction_rate;  // interned rate name per reaction
    std::vector<int64_t> ref_ptr{0};     // reaction -> refs
    std::vector<SpeciesRef> refs;
    std::vector<int32_t> species;        // declared order: interned names
    std::vector<int32_t> populations;
    std::vector<int32_t> species_pos;  // interned name -> declared index (-1: not declared)
};

enum Kind : uint8_t {
    kOther, kModel, kNumR, kNumS, kPList, kParam, kParamId, kParamExpr, kRList, kReaction, kRate, kReactants,
    kProducts, kSpRef, kSList, kSpecies, kSpId, kSpPop
};

bool captures(Kind k) {
    return k == kNumR || k == kNumS || k == kParamId || k == kParamExpr || k == kRate || k == kSpId || k == kSpPop;
}

class Walker {
public:
    explicit Walker(Doc &d) : d_(d) {}

    Kind classify(Kind parent, size_t depth, const std::string &n) {
        switch (parent) {
            case kOther:
                if (depth == 0 && !model_seen_ && n == "Model") {
                    model_seen_ = true;
                    return kModel;
                }
                return kOther;
            case kModel:
                if (n == "NumberOfReactions" && !d_.has_nr) return d_.has_nr = true, kNumR;
                if (n == "NumberOfSpecies" && !d_.has_ns) return d_.has_ns = true, kNumS;
                if (n == "ParametersList" && !d_.has_plist) return d_.has_plist = true, kPList;
                if (n == "ReactionsList" && !d_.has_rlist) return d_.has_rlist = true, kRList;
                if (n == "SpeciesList" && !d_.has_slist) return d_.has_slist = true, kSList;
                return kOther;
            case kPList:
                if (n == "Parameter") {
                    p_id_ = p_expr_ = false;
                    return kParam;
                }
                return kOther;
            case kParam:
                if (n == "Id" && !p_id_) return p_id_ = true, kParamId;
                if (n == "Expression" && !p_expr_) return p_expr_ = true, kParamExpr;
                return kOther;
            case kRList:
                if (n == "Reaction") {
                    r_rate_ = r_reac_ = r_prod_ = false;
                    r_first_ref_ = d_.refs.size();
                    return kReaction;
                }
                return kOther;
            case kReaction:
                if (n == "Rate" && !r_rate_) return r_rate_ = true, kRate;
                if (n == "Reactants" && !r_reac_) return r_reac_ = true, kReactants;
                if (n == "Products" && !r_prod_) return r_prod_ = true, kProducts;
                return kOther;
            case kReactants:
            case kProducts:
                return n == "SpeciesReference" ? kSpRef : kOther;
            case kSList:
                if (n == "Species") {
                    s_id_ = s_pop_ = false;
                    return kSpecies;
                }
                return kOther;
            case kSpecies:
                if (n == "Id" && !s_id_) return s_id_ = true, kSpId;
                if (n == "InitialPopulation" && !s_pop_) return s_pop_ = true, kSpPop;
                return kOther;
            default:
                return kOther;
        }
    }

    // a SpeciesReference start tag
    bool species_ref(const XmlPull &x, Kind parent, std::string &msg) {
        const std::string *id = x.attr("id"), *st = x.attr("stoichiometry");
        if (!id || !st) {
            msg = "SpeciesReference without id or stoichiometry";
            return false;
        }
        d_.refs.push_back({d_.species_names(*id), std::atoi(st->c_str()), parent == kReactants ? 0 : 1});
        return true;
    }

    bool close(Kind kind, std::string &text, std::string &msg) {
        switch (kind) {
            case kNumR: d_.n_reactions_declared = std::atoi(text.c_str()); return true;
            case kNumS: d_.n_species_declared = std::atoi(text.c_str()); return true;
            case kParamId: p_id_text_.swap(text); return true;
            case kParamExpr: p_expr_text_.swap(text); return true;
            case kParam: {
                if (!p_id_ || !p_expr_) {
                    msg = std::string("can not find ") + (!p_id_ ? "Id" : "Expression") + " node";
                    return false;
                }
                char *end = nullptr;
                const double v = std::strtod(p_expr_text_.c_str(), &end);
                if (end == p_expr_text_.c_str()) {
                    msg = "parameter " + p_id_text_ + ": expression '" + p_expr_text_ + "' is not a number";
                    return false;
                }
                const int32_t k = d_.rate_names(p_id_text_);
                if ((size_t)k >= d_.param_value.size()) {
                    d_.param_value.resize(k + 1, 0.0);
                    d_.param_set.resize(k + 1, 0);
                }
                d_.param_value[k] = v;  // a repeated Id takes the last value
                d_.param_set[k] = 1;
                d_.any_param = true;
                return true;
            }
            case kRate: r_rate_text_.swap(text); return true;
            case kReaction: {
                if (!r_rate_ || !r_reac_) {
                    msg = std::string("can not find ") + (!r_rate_ ? "Rate" : "Reactants") + " node";
                    return false;
                }
                // one entry per (side, species): a repeated reference keeps the last value
                size_t out = r_first_ref_;
                const size_t end = d_.refs.size();
                for (size_t i = r_first_ref_; i < end; ++i) {
                    bool later = false;
                    for (size_t k = i + 1; k < end; ++k)
                        later |= d_.refs[k].side == d_.refs[i].side && d_.refs[k].name == d_.refs[i].name;
                    if (!later) d_.refs[out++] = d_.refs[i];
                }
                d_.refs.resize(out);
                // files usually name reaction j's rate as parameter j: try that first
                const int32_t guess = (int32_t)d_.reaction_rate.size();
                d_.reaction_rate.push_back(d_.rate_names.is(guess, r_rate_text_) ? guess
                                                                                  : d_.rate_names(r_rate_text_));
                d_.ref_ptr.push_back((int64_t)d_.refs.size());
                d_.any_reaction = true;
                return true;
            }
            case kSpId: s_id_text_.swap(text); return true;
            case kSpPop: s_pop_text_.swap(text); return true;
            case kSpecies: {
                if (!s_id_ || !s_pop_) {
                    msg = std::string("can not find ") + (!s_id_ ? "Id" : "InitialPopulation") + " node";
                    return false;
                }
                const int32_t k = d_.species_names(s_id_text_);
                const int32_t pop = std::atoi(s_pop_text_.c_str());
                if ((size_t)k >= d_.species_pos.size()) d_.species_pos.resize(k + 1, -1);
                if (d_.species_pos[k] >= 0) {  // a repeated Id keeps its first index, takes the last population
                    d_.populations[d_.species_pos[k]] = pop;
                    return true;
                }
                d_.species_pos[k] = (int32_t)d_.species.size();
                d_.species.push_back(k);
                d_.populations.push_back(pop);
                return true;
            }
            default: return true;
        }
    }

private:
    Doc &d_;
    bool model_seen_ = false;
    bool p_id_ = false, p_expr_ = false, r_rate_ = false, r_reac_ = false, r_prod_ = false, s_id_ = false,
         s_pop_ = false;
    size_t r_first_ref_ = 0;
    std::string p_id_text_, p_expr_text_, r_rate_text_, s_id_text_, s_pop_text_;
};

tl_status read_doc(FILE *f, Doc &d, std::string &msg) {
    XmlPull x(f);
    Walker w(d);
    // open elements: kind, name and own text per depth (buffers reused)
    std::vector<Kind> kinds;
    std::vector<std::string> names, texts;
    size_t depth = 0;
    for (;;) {
        switch (x.next()) {
            case XmlPull::kError: msg = x.error(); return TL_ERR_PARSE;
            case XmlPull::kEof:
                if (depth > 0) {
                    msg = "missing </" + names[depth - 1] + ">";
                    return TL_ERR_PARSE;
                }
                return TL_OK;
            case XmlPull::kText:
                if (depth == 0) {
                    for (size_t i = 0; i < x.text_size(); ++i)
                        if (!std::isspace((unsigned char)x.text_data()[i])) {
                            msg = "unexpected character data outside an element";
                            return TL_ERR_PARSE;
                        }
                } else if (captures(kinds[depth - 1])) {
                    texts[depth - 1].append(x.text_data(), x.text_size());
                }
                break;
            case XmlPull::kStart: {
                const Kind parent = depth == 0 ? kOther : kinds[depth - 1];
                const Kind k = w.classify(parent, depth, x.name());
                if (k == kSpRef && !w.species_ref(x, parent, msg)) return TL_ERR_PARSE;
                if (depth == kinds.size()) {
                    kinds.push_back(k);
                    names.push_back(x.name());
                    texts.emplace_back();
                } else {
                    kinds[depth] = k;
                    names[depth].assign(x.name());
                    texts[depth].clear();
                }
                ++depth;
                break;
            }
            case XmlPull::kEnd: {
                if (depth == 0 || names[depth - 1] != x.name()) {
                    msg = "mismatched </" + x.name() + ">" +
                          (depth == 0 ? std::string() : " for <" + names[depth - 1] + ">");
                    return TL_ERR_PARSE;
                }
                --depth;
                std::string &t = texts[depth];
                if (captures(kinds[depth])) {
                    trim_in_place(t);
                    if (t.find('&') != std::string::npos) t = XmlPull::decode_entities(t);
                }
                if (!w.close(kinds[depth], t, msg)) return TL_ERR_PARSE;
                break;
            }
        }
    }
}

}  // namespace

// Read a StochKit file into m (N, M, CSC entries, rates, x0, names).
tl_status read_stochkit(FILE *f, tl_model &m, std::string &msg) {
    Doc d;
    const tl_status st = read_doc(f, d, msg);
    if (st != TL_OK) return st;
    auto missing = [&](const char *what) {
        msg = std::string("can not find ") + what + " node";
        return TL_ERR_PARSE;
    };
    if (!d.has_nr) return missing("NumberOfReactions");
    if (!d.has_ns) return missing("NumberOfSpecies");
    if (!d.has_plist) return missing("ParametersList");
    if (!d.any_param) return missing("Parameter");
    if (!d.has_rlist) return missing("ReactionsList");
    const int64_t M = (int64_t)d.reaction_rate.size();
    for (int64_t j = 0; j < M; ++j) {
        const int32_t k = d.reaction_rate[j];
        if ((size_t)k >= d.param_set.size() || !d.param_set[k]) {
            msg = "can not find the rate for the reaction " + d.rate_names.name(k);
            return TL_ERR_PARSE;
        }
    }
    if (!d.any_reaction) return missing("Reaction");
    if (M != d.n_reactions_declared) {
        msg = "reaction count mismatch: " + std::to_string(M) + " read, " + std::to_string(d.n_reactions_declared) +
              " declared";
        return TL_ERR_PARSE;
    }
    if (!d.has_slist) return missing("SpeciesList");
    if (d.species.empty()) return missing("Species");
    if ((int64_t)d.species.size() != d.n_species_declared) {
        msg = "species count mismatch: " + std::to_string(d.species.size()) + " read, " +
              std::to_string(d.n_species_declared) + " declared";
        return TL_ERR_PARSE;
    }

    // channel-major assembly: entries sorted by species index; a species in
    // both lists gives its reactant entry, then its product entry
    std::vector<int32_t> pos(d.species_names.size(), -1);
    for (size_t i = 0; i < d.species.size(); ++i) pos[d.species[i]] = (int32_t)i;
    m.N = (int32_t)d.species.size();
    m.M = (int32_t)M;
    m.col_ptr.assign(1, 0);
    m.row_ind.clear();
    m.stoich.clear();
    m.rates.resize(M);
    m.row_ind.reserve(d.refs.size());
    m.stoich.reserve(d.refs.size());
    m.col_ptr.reserve(M + 1);
    std::vector<std::pair<int64_t, int32_t>> ent;  // (species * 2 + side, value)
    for (int64_t j = 0; j < M; ++j) {
        ent.clear();
        for (int64_t r = d.ref_ptr[j]; r < d.ref_ptr[j + 1]; ++r) {
            const SpeciesRef &s = d.refs[r];
            const int32_t i = pos[s.name];
            if (i < 0) {
                msg = "reaction uses undeclared species " + d.species_names.name(s.name);
                return TL_ERR_PARSE;
            }
            ent.push_back({(int64_t)i * 2 + s.side, s.side == 0 ? -s.v : s.v});
        }
        std::sort(ent.begin(), ent.end());
        for (auto &e : ent) {
            if (e.second < -128 || e.second > 127) {
                msg = "stoichiometry out of range";
                return TL_ERR_STOICH_RANGE;
            }
            m.row_ind.push_back((int32_t)(e.first >> 1));
            m.stoich.push_back((int8_t)e.second);
        }
        m.col_ptr.push_back((int32_t)m.row_ind.size());
        m.rates[j] = d.param_value[d.reaction_rate[j]];
    }
    m.x0 = d.populations;
    m.names.resize(m.N);
    for (int32_t i = 0; i < m.N; ++i) m.names[i] = d.species_names.name(d.species[i]);
    return TL_OK;
}

}  // namespace tl
