// toy_system_gen -- random StochKit 2.0 model generator (XML on stdout).
//
// Same families and options as the reference generator
// (toy_system_gen/main.cpp:240-876):
//   -r <species> [-c <channels>] [-t1 n1 -t2 n2 -t3 n3]   random network of
//        type 1 (S -> products), type 2 (Sa + Sb -> products) and type 3
//        (2S -> S') channels, every species used at least once (main.cpp:562-692)
//   -w   weakly coupled cyclic chain S_i -> S_(i+1), channels = species (main.cpp:694-710)
//   -s   colloidal aggregation / fragmentation, channels = species^2/2 (main.cpp:712-756)
//   -mnp / -mxp   initial populations uniform in [mnp, mxp] (default 0..10000)
// Extensions: --seed N (the reference seeds with time()), --toy (1-2
// distinct reactants with coefficient 1 and 1-2 distinct products, the
// benchmark recipe), --log-rates lo hi (log-uniform rates printed with full
// precision; the default is the reference's uniform [0, 1] printed "%f").
// Fixes: random species indices are drawn in [1, species] (the reference
// draws in [1, channels], main.cpp:551,673); channels come out in generation
// order (the reference's order is that of an unordered_set).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Rx {
    int type = 1;  // 1, 2, 3 (for -w/-s the reactant pattern decides)
    std::vector<std::pair<int, int>> reac, prod;  // (species 1-based, stoichiometry)
};

struct Opt {
    int species = 0, channels = 0, t1 = -1, t2 = -1, t3 = -1;
    int mnp = 0, mxp = 10000;
    char family = 'n';  // n, w, s, t (toy)
    bool log_rates = false;
    double rate_lo = 1e-3, rate_hi = 1.0;
    unsigned long long seed = 0;
    bool seed_set = false;
};

[[noreturn]] void die(const std::string &m) {
    std::cerr << m << std::endl;
    std::exit(EXIT_FAILURE);
}

int next_int(std::mt19937_64 &g, int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(g); }

// Removals from the unused list of the channel being drawn, (position, species),
// so a rejected channel can put them back without copying the list.
std::vector<std::pair<int, int>> undo_log;

void undo_removals(std::vector<int> &unused) {
    for (auto it = undo_log.rbegin(); it != undo_log.rend(); ++it) {
        const int i = it->first, s = it->second;
        if (i == (int)unused.size()) {
            unused.push_back(s);
        } else {
            unused.push_back(unused[i]);
            unused[i] = s;
        }
    }
    undo_log.clear();
}

// A species not used yet if any remain, else any species; never one of `avoid`.
int next_species(std::mt19937_64 &g, std::vector<int> &unused, int n, const std::set<int> &avoid) {
    for (int attempt = 0; attempt < 1000; ++attempt) {
        if (!unused.empty()) {
            int i = next_int(g, 0, (int)unused.size() - 1);
            int s = unused[i];
            if (avoid.count(s)) {
                bool any = false;
                for (int u : unused) any |= !avoid.count(u);
                if (!any) goto fallback;
                continue;
            }
            undo_log.push_back({i, s});
            unused[i] = unused.back();
            unused.pop_back();
            return s;
        }
    fallback:
        int s = next_int(g, 1, n);
        if (!avoid.count(s)) return s;
    }
    die("There are not enough reactants to produce unique reactions");
}

std::string key_of(const Rx &r) {  // identity used for uniqueness (Reaction::operator==)
    std::ostringstream o;
    std::vector<int> rs;
    for (auto &p : r.reac) rs.push_back(p.first * 8 + p.second);
    std::sort(rs.begin(), rs.end());
    o << r.type << ":";
    for (int v : rs) o << v << ",";
    if (r.type != 2 && !r.prod.empty()) o << "p" << r.prod[0].first;
    return o.str();
}

std::vector<Rx> gen_normal(std::mt19937_64 &g, const Opt &o) {
    std::vector<int> unused;
    for (int i = 1; i <= o.species; ++i) unused.push_back(i);
    std::vector<Rx> out;
    std::set<std::string> seen;
    const int counts[3] = {o.t1, o.t2, o.t3};
    for (int type = 1; type <= 3; ++type)
        for (int n = 0; n < counts[type - 1]; ++n) {
            int tries = 0;
            for (;; ++tries) {
                if (tries >= 100) die("There are not enough reactants to produce unique reactions");
                undo_log.clear();
                Rx r;
                r.type = type;
                std::set<int> used;
                if (type == 1) {
                    int a = next_species(g, unused, o.species, used);
                    used.insert(a);
                    r.reac = {{a, 1}};
                    int ps = next_int(g, 0, 2);
                    if (ps) {
                        int b = next_species(g, unused, o.species, used);
                        used.insert(b);
                        r.prod.push_back({b, ps});
                        if (ps == 1 && next_int(g, 0, 1)) {
                            int c = next_species(g, unused, o.species, used);
                            r.prod.push_back({c, 1});
                        }
                    }
                } else if (type == 2) {
                    int a = next_species(g, unused, o.species, used);
                    used.insert(a);
                    int b = next_species(g, unused, o.species, used);
                    used.insert(b);
                    r.reac = {{a, 1}, {b, 1}};
                    int c = next_species(g, unused, o.species, used);
                    used.insert(c);
                    r.prod.push_back({c, 1});
                    if (next_int(g, 0, 1)) r.prod.push_back({next_species(g, unused, o.species, used), 1});
                } else {
                    int a = next_species(g, unused, o.species, used);
                    used.insert(a);
                    r.reac = {{a, 2}};
                    r.prod.push_back({next_species(g, unused, o.species, used), 1});
                }
                if (seen.insert(key_of(r)).second) {
                    out.push_back(r);
                    break;
                }
                undo_removals(unused);
            }
        }
    if (!unused.empty())
        die("Not all reactants were used, " + std::to_string(unused.size()) + " left. Too few channels(reactions)");
    return out;
}

// 1-2 distinct reactants and 1-2 distinct products, all different species:
// the first nr + np species of a uniform random order, drawn one at a time by
// rejection (O(1) per channel, so 1e6 channels over 1e5 species stay fast).
std::vector<Rx> gen_toy(std::mt19937_64 &g, const Opt &o) {
    std::vector<Rx> out;
    out.reserve(o.channels);
    for (int n = 0; n < o.channels; ++n) {
        Rx r;
        int nr = next_int(g, 1, 2), np = next_int(g, 1, 2);
        int pick[4];
        for (int i = 0; i < nr + np; ++i) {
            int s;
            do {
                s = next_int(g, 1, o.species);
            } while (std::find(pick, pick + i, s) != pick + i);
            pick[i] = s;
        }
        for (int i = 0; i < nr; ++i) r.reac.push_back({pick[i], 1});
        for (int i = 0; i < np; ++i) r.prod.push_back({pick[nr + i], 1});
        r.type = nr == 2 ? 2 : 1;
        out.push_back(r);
    }
    return out;
}

std::vector<Rx> gen_cyclic(const Opt &o) {
    std::vector<Rx> out;
    for (int i = 1; i <= o.species; ++i) {
        Rx r;
        r.reac = {{i, 1}};
        r.prod = {{i < o.species ? i + 1 : 1, 1}};
        out.push_back(r);
    }
    return out;
}

std::vector<Rx> gen_colloid(const Opt &o) {
    std::vector<Rx> out;
    const int N = o.species;
    for (int n = 1; n <= N / 2; ++n) {
        Rx d;
        d.reac = {{n, 2}};
        d.prod = {{2 * n, 1}};
        d.type = 3;
        out.push_back(d);
        for (int m = n + 1; m <= N - n; ++m) {
            Rx r;
            r.reac = {{n, 1}, {m, 1}};
            r.prod = {{m + n, 1}};
            r.type = 2;
            out.push_back(r);
        }
    }
    for (int p = 1; p <= N; ++p)
        for (int q = 1; q <= p / 2; ++q) {
            Rx r;
            r.reac = {{p, 1}};
            if (q == p - q) r.prod = {{q, 2}};
            else r.prod = {{q, 1}, {p - q, 1}};
            out.push_back(r);
        }
    return out;
}

std::string desc(const Rx &r) {
    std::ostringstream o;
    for (size_t i = 0; i < r.reac.size(); ++i)
        o << (i ? " + " : "") << (r.reac[i].second == 2 ? "2 * " : "") << "S" << r.reac[i].first;
    o << " to";
    for (size_t i = 0; i < r.prod.size(); ++i)
        o << (i ? " +" : "") << " " << (r.prod[i].second == 2 ? "2 * " : "") << "S" << r.prod[i].first;
    if (r.prod.empty()) o << " nothing";
    return o.str();
}

}  // namespace

int main(int argc, char **argv) {
    std::ios::sync_with_stdio(false);
    Opt o;
    for (int i = 1; i < argc; ++i) {
        const std::string k = argv[i];
        auto val = [&]() -> const char * {
            if (i + 1 >= argc) die("Use -c <channels number> -r <reactants number> options at least");
            return argv[++i];
        };
        if (k == "-c") o.channels = atoi(val());
        else if (k == "-r") o.species = atoi(val());
        else if (k == "-t1") o.t1 = atoi(val());
        else if (k == "-t2") o.t2 = atoi(val());
        else if (k == "-t3") o.t3 = atoi(val());
        else if (k == "-mnp") o.mnp = atoi(val());
        else if (k == "-mxp") o.mxp = atoi(val());
        else if (k == "-w") o.family = 'w';
        else if (k == "-s") o.family = 's';
        else if (k == "--toy") o.family = 't';
        else if (k == "--seed") {
            o.seed = strtoull(val(), nullptr, 10);
            o.seed_set = true;
        } else if (k == "--log-rates") {
            o.log_rates = true;
            o.rate_lo = atof(val());
            o.rate_hi = atof(val());
        } else die("Unknown parameter: " + k);
    }
    if (o.species <= 3) die("Incorrect number of reactants");
    if (o.mnp < 0 || o.mxp < o.mnp) die("Incorrect population range");
    std::mt19937_64 g(o.seed_set ? o.seed : (unsigned long long)time(nullptr));
    if (o.family == 'w') o.channels = o.species;
    if (o.family == 's') o.channels = o.species * o.species / 2;
    if (o.channels <= 0) die("Incorrect number of channels");
    if (o.family == 'n') {
        // unspecified type counts share what the specified ones leave
        int given = (o.t1 > 0 ? o.t1 : 0) + (o.t2 > 0 ? o.t2 : 0) + (o.t3 > 0 ? o.t3 : 0);
        if (given > o.channels) die("Inconsistent number of reactions of different types");
        int left = o.channels - given;
        if (o.t1 < 0) {
            o.t1 = (o.t2 >= 0 && o.t3 >= 0) ? left : next_int(g, 0, left);
            left -= o.t1;
        }
        if (o.t2 < 0) {
            o.t2 = (o.t3 >= 0) ? left : next_int(g, 0, left);
            left -= o.t2;
        }
        if (o.t3 < 0) o.t3 = left;
        if (o.t1 + o.t2 + o.t3 != o.channels) die("Inconsistent number of reactions of different types");
    }
    std::vector<Rx> rx;
    switch (o.family) {
        case 'w': rx = gen_cyclic(o); break;
        case 's': rx = gen_colloid(o); break;
        case 't': rx = gen_toy(g, o); break;
        default: rx = gen_normal(g, o); break;
    }
    const int M = (int)rx.size();
    std::cerr << "generating " << M << " reactions over " << o.species << " species" << std::endl;
    // written straight to stdout (1e6-channel models are about half a
    // gigabyte of text); cout is unsynchronised from stdio, so it buffers
    std::ostream &x = std::cout;
    x << "<!-- Stochkit 2.0 text input format\n\nreference:\nM. Rathinam, L.R. Petzold, Y. Cao, and D.T. "
         "Gillespie. J Chem Phys 119, 12784 (2003).\n-->\n\n";
    x << "<!-- generated by toy_system_gen (tau_leaping_amd)";
    if (o.seed_set) x << ", seed " << o.seed;
    x << " -->\n";
    if (o.family == 'w') x << "<!--Weakly coupled (Cyclic Chain) system -->\n";
    if (o.family == 's') x << "<!--Strongly coupled (Colloidal Aggregation) system -->\n";
    x << "<Model>\n\t<Description>Test input data</Description>\n";
    x << "\t<NumberOfReactions>" << M << "</NumberOfReactions>\n\t<NumberOfSpecies>" << o.species
      << "</NumberOfSpecies>\n\t<ParametersList>\n";
    char buf[64];
    for (int j = 0; j < M; ++j) {
        double c;
        if (o.log_rates)
            c = std::exp(std::uniform_real_distribution<double>(std::log(o.rate_lo), std::log(o.rate_hi))(g));
        else
            c = std::uniform_real_distribution<double>(0.0, 1.0)(g);
        if (o.log_rates) snprintf(buf, sizeof buf, "%.17g", c);
        else snprintf(buf, sizeof buf, "%f", c);
        x << "\t\t<Parameter>\n\t\t\t<Id>c" << j + 1 << "</Id>\n\t\t\t<Expression>" << buf
          << "</Expression>\n\t\t</Parameter>\n";
    }
    x << "\t</ParametersList>\n\t<ReactionsList>\n";
    for (int j = 0; j < M; ++j) {
        x << "\t\t<Reaction>\n\t\t\t<Id>R" << j + 1 << "</Id>\n\t\t\t<Description>" << desc(rx[j])
          << "</Description>\n\t\t\t<Type>mass-action</Type>\n\t\t\t<Rate>c" << j + 1 << "</Rate>\n\t\t\t<Reactants>\n";
        for (auto &p : rx[j].reac)
            x << "\t\t\t\t<SpeciesReference id=\"S" << p.first << "\" stoichiometry=\"" << p.second << "\"/>\n";
        x << "\t\t\t</Reactants>\n\t\t\t<Products>\n";
        for (auto &p : rx[j].prod)
            x << "\t\t\t\t<SpeciesReference id=\"S" << p.first << "\" stoichiometry=\"" << p.second << "\"/>\n";
        x << "\t\t\t</Products>\n\t\t</Reaction>\n";
    }
    x << "\t</ReactionsList>\n\t<SpeciesList>\n";
    for (int i = 1; i <= o.species; ++i)
        x << "\t\t<Species>\n\t\t\t<Id>S" << i << "</Id>\n\t\t\t<Description>Species #" << i
          << "</Description>\n\t\t\t<InitialPopulation>" << next_int(g, o.mnp, o.mxp)
          << "</InitialPopulation>\n\t\t</Species>\n";
    x << "\t</SpeciesList>\n</Model>\n";
    x.flush();
    std::cerr << "Finished successfully" << std::endl;
    return EXIT_SUCCESS;
}
