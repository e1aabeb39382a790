// qg_vps.cpp — reader of MBK norm-conserving pseudopotential files (.vps) and their
// conversion to the engine's tabulated projector channels.
//
// Reference semantics (QUMASUN source/):
//   file model: key/value lines and "<name" ... "name>" blocks, '*' and '#' start a comment,
//   "<VPS>" starts a comment on key lines, "/*" ... "*/" lines are skipped, the first
//   occurrence of a key or block wins                              JobParam.h:246-384
//   LoadVPS: vps.type MBK, pseudo.NandL cutoffs per l, project.energies, Pseudo.Potentials
//   (xi, r, V_local, two j-components per projector), density.PCC, valence.electron,
//   local.cutoff, SortByL                                          vps_loader.h:87-247
//   projector channel k of an atom: l, E = energy_up, cutoff = cutoff_r[l], radial values the
//   mean of the two j-components, r < radius[0] takes the first value
//                                                                  PseudoPotOperator_mpi_v2.h:64-84,
//                                                                  :108-166
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include "qg_internal.hpp"

struct qg_vps {
    int valence_electron = 0;
    int num_projectors = 0;
    int num_radial_grids = 0;
    double cutoff_vlocal = 0.0;
    double cutoff_r[4] = {0.0, 0.0, 0.0, 0.0};
    double xi_min = 0.0, xi_delta = 0.0;
    int has_pcc = 0;
    double pcc_integral = 0.0;
    std::vector<int> l;
    std::vector<double> e_up, e_down;
    std::vector<double> radius, v_local, pcc;
    std::vector<std::vector<double>> comp;   // 2 per projector: the two j-components
    std::vector<int> comp_of;                // projector slot 2k + s -> its table in comp
    std::vector<std::vector<double>> mean;   // per projector: the mean of its two components
};

namespace qg {
namespace {

constexpr const char* kDelim = " \t";

bool in_set(char c, const char* set) { return c != '\0' && std::strchr(set, c) != nullptr; }

// first character not in set (nullptr at the end of the string)
char* skip_set(char* p, const char* set) {
    for (; *p; ++p)
        if (!in_set(*p, set)) return p;
    return nullptr;
}
// first character in set (nullptr if none)
char* find_set(char* p, const char* set) {
    for (; *p; ++p)
        if (in_set(*p, set)) return p;
    return nullptr;
}
void rtrim(char* s) {
    for (size_t n = std::strlen(s); n > 0 && in_set(s[n - 1], kDelim); --n) s[n - 1] = '\0';
}

// The job-file model of the reader: values of key lines and the lines of blocks, for the MBK
// conventions (block start "<name", block end "name>", comment characters "*#\r\n", comment
// word "<VPS>").  Lines are read 1023 characters at a time, as the reference reads them.
struct JobText {
    std::map<std::string, std::string> values;
    std::map<std::string, std::vector<std::string>> blocks;

    bool parse(const char* path) {
        FILE* fp = std::fopen(path, "r");
        if (!fp) return false;
        const char* kComment = "*#\r\n";
        const char* kWord = "<VPS>";
        char line[1024];
        bool in_comment = false;
        auto skip_comment_lines = [&](const char* s) {
            if (in_comment) {
                if (std::strncmp(s, "*/", 2) == 0) in_comment = false;
                return true;
            }
            if (std::strncmp(s, "/*", 2) == 0) {
                in_comment = true;
                return true;
            }
            return false;
        };
        while (std::fgets(line, sizeof(line), fp)) {
            if (skip_comment_lines(line)) continue;
            if (char* w = std::strstr(line, kWord)) *w = '\0';
            rtrim(line);
            if (char* c = find_set(line, kComment)) *c = '\0';
            rtrim(line);
            char* key = skip_set(line, kDelim);
            if (!key) continue;
            char* value = nullptr;
            if (char* ke = find_set(key + 1, kDelim)) {
                *ke = '\0';
                value = skip_set(ke + 1, kDelim);
            }
            if (key[0] != '<') {
                if (value) values.emplace(key, value);
                continue;
            }
            // block: lines up to "name>"
            const std::string name(key + 1);
            const std::string term = name + ">";
            std::vector<std::string> lines;
            bool closed = false;
            while (std::fgets(line, sizeof(line), fp)) {
                if (skip_comment_lines(line)) continue;
                if (char* c = find_set(line, kComment)) *c = '\0';
                rtrim(line);
                char* head = skip_set(line, kDelim);
                if (!head) continue;
                if (term == head) {
                    closed = true;
                    break;
                }
                lines.emplace_back(head);
            }
            if (closed) blocks.emplace(name, std::move(lines));
        }
        std::fclose(fp);
        return true;
    }
    const std::vector<std::string>* block(const char* k) const {
        auto it = blocks.find(k);
        return it == blocks.end() ? nullptr : &it->second;
    }
    const std::string* value(const char* k) const {
        auto it = values.find(k);
        return it == values.end() ? nullptr : &it->second;
    }
};

bool iequal(const std::string& a, const char* b) {
    if (a.size() != std::strlen(b)) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
    return true;
}

// numbers of one line read left to right as strtol / strtod read them (a missing number is 0)
struct Fields {
    const char* p;
    explicit Fields(const std::string& s) : p(s.c_str()) {}
    long next_int() {
        char* e = nullptr;
        const long v = std::strtol(p, &e, 10);
        p = e;
        return v;
    }
    double next_double() {
        char* e = nullptr;
        const double v = std::strtod(p, &e);
        p = e;
        return v;
    }
};

void load_vps(const char* path, qg_vps& pp) {
    JobText job;
    QG_REQUIRE(job.parse(path), std::string("cannot open ") + path);
    const std::string* type = job.value("vps.type");
    QG_REQUIRE(type && iequal(*type, "MBK"), "VPS of MBK is only supported (vps.type)");

    if (const auto* b = job.block("pseudo.NandL")) {
        for (const std::string& line : *b) {
            Fields f(line);
            f.next_int();                       // index
            f.next_int();                       // principal quantum number
            const long l = f.next_int();
            const double cutoff = f.next_double();
            QG_REQUIRE(l >= 0, "pseudo.NandL: negative l");
            if (l < 4 && pp.cutoff_r[l] < cutoff) pp.cutoff_r[l] = cutoff;
        }
    }

    const auto* en = job.block("project.energies");
    QG_REQUIRE(en && !en->empty(), "project.energies block missing");
    {
        Fields f((*en)[0]);
        const long num = f.next_int();
        QG_REQUIRE(num >= 0 && num < 4096, "project.energies: bad projector count");
        QG_REQUIRE((long)en->size() - 1 == num, "project.energies: the count differs from the number of lines");
        pp.num_projectors = (int)num;
        for (long i = 1; i <= num; ++i) {
            Fields g((*en)[i]);
            const long l = g.next_int();
            QG_REQUIRE(l >= 0 && l < 4, "project.energies: l out of range");
            pp.l.push_back((int)l);
            pp.e_up.push_back(g.next_double());
            pp.e_down.push_back(g.next_double());
        }
    }

    const auto* pot = job.block("Pseudo.Potentials");
    QG_REQUIRE(pot && pot->size() >= 2, "Pseudo.Potentials block missing or shorter than two lines");
    const int ng = (int)pot->size();
    pp.num_radial_grids = ng;
    pp.radius.resize(ng);
    pp.v_local.resize(ng);
    pp.comp.assign(2 * (size_t)pp.num_projectors, std::vector<double>(ng));
    double xi = 0.0;
    for (int i = 0; i < ng; ++i) {
        Fields f((*pot)[i]);
        xi = f.next_double();
        if (i == 0) pp.xi_min = xi;
        pp.radius[i] = f.next_double();
        pp.v_local[i] = f.next_double();
        for (int k = 0; k < 2 * pp.num_projectors; ++k) pp.comp[k][i] = f.next_double();
    }
    pp.xi_delta = (xi - pp.xi_min) / (double)(ng - 1);

    if (const auto* pc = job.block("density.PCC"); pc && !pc->empty()) {
        QG_REQUIRE((int)pc->size() == ng, "density.PCC has incorrect grid size");
        pp.pcc.resize(ng);
        double sum = 0.0;
        for (int i = 0; i < ng; ++i) {
            Fields f((*pc)[i]);
            f.next_double();                    // xi
            const double r = f.next_double();
            pp.pcc[i] = f.next_double();
            sum += pp.pcc[i] * r * r * r;
        }
        pp.has_pcc = 1;
        pp.pcc_integral = sum * pp.xi_delta * 4.0 * M_PI;
    }
    const std::string* ve = job.value("valence.electron");
    pp.valence_electron = ve ? (int)std::strtod(ve->c_str(), nullptr) : 0;
    const std::string* lc = job.value("local.cutoff");
    pp.cutoff_vlocal = lc ? std::strtod(lc->c_str(), nullptr) : 0.0;

    // PseudoPot_MBK::SortByL (vps_loader.h:67-82) as it behaves: for a projector whose l is
    // below its predecessor's (in the pass's index range), the predecessor's l, energies and
    // tables are copied over it.  Files ordered by l pass unchanged.
    const int n = pp.num_projectors;
    pp.comp_of.resize(2 * (size_t)n);
    for (int k = 0; k < 2 * n; ++k) pp.comp_of[k] = k;
    for (int k = 1; k < n; ++k) {
        for (int j = 2; j < n - j; ++j) {
            if (pp.l[j] < pp.l[j - 1]) {
                pp.l[j] = pp.l[j - 1];
                pp.e_up[j] = pp.e_up[j - 1];
                pp.e_down[j] = pp.e_down[j - 1];
                pp.comp_of[2 * j] = pp.comp_of[2 * (j - 1)];
                pp.comp_of[2 * j + 1] = pp.comp_of[2 * (j - 1) + 1];
            }
        }
    }
    pp.mean.assign(n, std::vector<double>(ng));
    for (int k = 0; k < n; ++k) {
        const auto& a = pp.comp[pp.comp_of[2 * k]];
        const auto& b = pp.comp[pp.comp_of[2 * k + 1]];
        for (int i = 0; i < ng; ++i) pp.mean[k][i] = (a[i] + b[i]) / 2.0;
    }
}

}  // namespace
}  // namespace qg

using namespace qg;

extern "C" {

int qg_vps_load(const char* path, qg_vps** out) {
    return guarded([&] {
        QG_REQUIRE(path && out, "null argument");
        *out = nullptr;
        auto* v = new qg_vps;
        try {
            load_vps(path, *v);
        } catch (...) {
            delete v;
            throw;
        }
        *out = v;
    });
}

int qg_vps_destroy(qg_vps* v) {
    delete v;
    return QG_OK;
}

int qg_vps_get_info(const qg_vps* v, qg_vps_info* info) {
    return guarded([&] {
        QG_REQUIRE(v && info, "null argument");
        info->valence_electron = v->valence_electron;
        info->num_projectors = v->num_projectors;
        info->num_radial_grids = v->num_radial_grids;
        info->cutoff_vlocal = v->cutoff_vlocal;
        for (int l = 0; l < 4; ++l) info->cutoff_r[l] = v->cutoff_r[l];
        info->xi_min = v->xi_min;
        info->xi_delta = v->xi_delta;
        info->has_pcc_charge = v->has_pcc;
        info->pcc_charge_integral = v->pcc_integral;
        int tot = 0, maxl = 0;
        for (int k = 0; k < v->num_projectors; ++k) {
            tot += 2 * v->l[k] + 1;
            maxl = std::max(maxl, v->l[k]);
        }
        info->total_projector_lm = tot;
        info->max_projector_l = maxl;
    });
}

int qg_vps_get_radial(const qg_vps* v, double* radius, double* v_local, double* pcc_charge) {
    return guarded([&] {
        QG_REQUIRE(v, "null argument");
        const size_t n = (size_t)v->num_radial_grids;
        if (radius) std::memcpy(radius, v->radius.data(), n * sizeof(double));
        if (v_local) std::memcpy(v_local, v->v_local.data(), n * sizeof(double));
        if (pcc_charge) {
            if (v->has_pcc) std::memcpy(pcc_charge, v->pcc.data(), n * sizeof(double));
            else std::memset(pcc_charge, 0, n * sizeof(double));
        }
    });
}

int qg_vps_get_projector(const qg_vps* v, int k, int* l, double* energy_up, double* energy_down, double* up,
                         double* down) {
    return guarded([&] {
        QG_REQUIRE(v && k >= 0 && k < v->num_projectors, "projector index out of range");
        const size_t n = (size_t)v->num_radial_grids;
        if (l) *l = v->l[k];
        if (energy_up) *energy_up = v->e_up[k];
        if (energy_down) *energy_down = v->e_down[k];
        if (up) std::memcpy(up, v->comp[v->comp_of[2 * k]].data(), n * sizeof(double));
        if (down) std::memcpy(down, v->comp[v->comp_of[2 * k + 1]].data(), n * sizeof(double));
    });
}

int qg_vps_projector_specs(const qg_vps* v, qg_proj_spec* specs, int cap) {
    return guarded([&] {
        QG_REQUIRE(v && (specs || cap == 0), "null argument");
        QG_REQUIRE(cap >= v->num_projectors, "spec array shorter than the projector count");
        for (int k = 0; k < v->num_projectors; ++k) {
            qg_proj_spec& s = specs[k];
            std::memset(&s, 0, sizeof(s));
            s.l = v->l[k];
            s.energy = v->e_up[k];
            s.cutoff = v->cutoff_r[v->l[k]];
            s.radial_kind = QG_RADIAL_TABLE;
            s.n_table = v->num_radial_grids;
            s.xi_min = v->xi_min;
            s.xi_delta = v->xi_delta;
            s.r_min = v->radius[0];
            s.table = v->mean[k].data();
        }
    });
}

}  // extern "C"
