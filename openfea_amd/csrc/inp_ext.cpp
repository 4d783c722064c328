// inp_ext.cpp -- keyword-level reading of the Abaqus deck features that the
// reference's reader does not take (it reads one *Cload line, the node sets
// "Set-load"/"Set-fix" and one global material: src/data2cae.cpp:315-446).
//
// Parsed blocks: *Nset / *Elset (explicit ids, set names, "generate"),
// *Surface type=ELEMENT ("elset_or_element, S#"), every *Cload line,
// *Dsload (P pressure and TRVEC traction), *Boundary (ENCASTRE, PINNED,
// XSYMM/YSYMM/ZSYMM, "first[, last[, value]]"), *Initial Conditions
// type=VELOCITY, *Material with *Elastic / *Density and *Solid Section
// (elset -> material).  Node and element ids map to line-order indices
// id - 1, the reference's convention (src/data2cae.cpp:92-158).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "ofea_internal.h"

namespace ofea {
namespace {

std::string lower(std::string s)
{
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    std::string t = s.substr(a, b - a);
    if (t.size() >= 2 && t.front() == '"' && t.back() == '"') t = t.substr(1, t.size() - 2);
    return t;
}

std::vector<std::string> split_csv(const std::string &line)
{
    std::vector<std::string> out;
    std::string tok;
    std::istringstream iss(line);
    while (std::getline(iss, tok, ',')) out.push_back(trim(tok));
    while (!out.empty() && out.back().empty()) out.pop_back();
    return out;
}

bool to_int(const std::string &s, long &v)
{
    if (s.empty()) return false;
    char *end = nullptr;
    v = std::strtol(s.c_str(), &end, 10);
    return *end == '\0';
}

bool to_double(const std::string &s, double &v)
{
    if (s.empty()) return false;
    char *end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return *end == '\0';
}

struct Keyword {
    std::string name;                           // lower case, without '*'
    std::map<std::string, std::string> params;  // lower-case keys; values as written ("" for flags)
    std::vector<std::vector<std::string>> data;
};

// Reads the deck as keyword blocks ("**" comment lines dropped, CR/LF safe).
int read_keywords(const std::string &path, std::vector<Keyword> &out)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(OFEA_ERR_IO, "cannot open " + path);
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.size() >= 2 && line[0] == '*' && line[1] == '*') continue;
        if (!line.empty() && line[0] == '*') {
            Keyword k;
            auto parts = split_csv(line.substr(1));
            if (parts.empty()) continue;
            k.name = lower(parts[0]);
            for (size_t i = 1; i < parts.size(); ++i) {
                size_t eq = parts[i].find('=');
                if (eq == std::string::npos)
                    k.params[lower(trim(parts[i]))] = "";
                else
                    k.params[lower(trim(parts[i].substr(0, eq)))] = trim(parts[i].substr(eq + 1));
            }
            out.push_back(std::move(k));
        } else if (!out.empty() && !trim(line).empty()) {
            out.back().data.push_back(split_csv(line));
        }
    }
    return OFEA_OK;
}

// Face node lists (0-based local nodes) of the Abaqus faces S1..S6 (C3D8) and
// S1..S4 (C3D4).
const int HEX_FACES[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};
const int TET_FACES[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {2, 3, 0}};

struct Cross {
    double x, y, z;
};
Cross sub(const double *a, const double *b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
Cross cross(Cross a, Cross b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

}  // namespace

// Abaqus part/instance scoping.  Sets defined between *Part and *End Part
// belong to that part and are stored under "part.set" (and, for decks that
// name them unqualified, under "set" as well); *Instance, name=I, part=P lets
// the assembly and the steps name them "I.set" (the reference's own decks,
// e.g. sample_model/NLFEA.inp:82956-82978, write "Part-1-1.Set-fix").  Ids
// written inside a part, or qualified by an instance, are labels of that
// part: with the reference's line-order numbering, node L of a part is
// global index (nodes of the earlier parts) + L - 1, and elements likewise.
struct Scope {
    std::string part;  // lower case; empty outside *Part ... *End Part
    int node_off = 0, elem_off = 0;
};
struct PartInfo {
    int node_off = 0, elem_off = 0;
};

// Resolves a "node id or node set" target to 0-based node indices.
static int resolve(const std::map<std::string, std::vector<int>> &sets, const std::string &tok, int count,
                   std::vector<int> &out, const char *what, const Scope &sc = Scope(),
                   const std::map<std::string, std::string> *inst = nullptr,
                   const std::map<std::string, PartInfo> *parts = nullptr, bool nodes = true)
{
    out.clear();
    long id;
    auto by_id = [&](long lab, int off) -> int {
        const long g = lab + off;
        if (lab < 1 || g > count) return fail(OFEA_ERR_IO, std::string(what) + " id out of range: " + tok);
        out.push_back((int)g - 1);
        return OFEA_OK;
    };
    if (to_int(tok, id)) return by_id(id, nodes ? sc.node_off : sc.elem_off);
    const std::string t = lower(tok);
    // instance-qualified: "instance.set" or "instance.label"
    const size_t dot = t.find('.');
    if (dot != std::string::npos && inst) {
        auto ii = inst->find(t.substr(0, dot));
        if (ii != inst->end()) {
            const std::string rest = t.substr(dot + 1);
            int off = 0;
            if (parts) {
                auto pi = parts->find(ii->second);
                if (pi != parts->end()) off = nodes ? pi->second.node_off : pi->second.elem_off;
            }
            if (to_int(rest, id)) return by_id(id, off);
            auto it = sets.find(ii->second + "." + rest);
            if (it != sets.end()) {
                out = it->second;
                return OFEA_OK;
            }
        }
    }
    if (!sc.part.empty()) {  // inside a part its own sets come first
        auto it = sets.find(sc.part + "." + t);
        if (it != sets.end()) {
            out = it->second;
            return OFEA_OK;
        }
    }
    auto it = sets.find(t);
    if (it == sets.end()) return fail(OFEA_ERR_IO, std::string("unknown ") + what + " set: " + tok);
    out = it->second;
    return OFEA_OK;
}

int parse_inp_ext(const std::string &path, ofea_deck &D)
{
    D.ext_error.clear();
    D.sections.clear();
    std::vector<Keyword> kws;
    int rc = read_keywords(path, kws);
    if (rc) return rc;
    const int nd = D.nd, ne = D.ne;
    std::map<std::string, std::vector<int>> nsets, elsets;
    std::map<std::string, std::vector<std::pair<int, int>>> surfaces;  // (element, face 0-based)
    struct Mat {
        double E = NAN, nu = NAN, rho = NAN;
    };
    std::map<std::string, Mat> mats;
    std::string cur_mat;
    D.f_nodal.assign(3 * (size_t)nd, 0.0);
    D.v0.assign(3 * (size_t)nd, 0.0);
    D.bc_mask.assign((size_t)nd, 0);
    D.bc_nonzero = 0;
    D.E_elem.assign((size_t)ne, NAN);
    D.nu_elem.assign((size_t)ne, NAN);
    D.rho_elem.assign((size_t)ne, NAN);
    std::vector<int> tgt;
    Scope scope;                                // current part (empty at model/assembly level)
    std::map<std::string, PartInfo> parts;      // part -> offsets of its nodes and elements
    std::map<std::string, std::string> inst;    // instance -> part
    int node_rows = 0, elem_rows = 0;           // *Node / *Element data lines read so far
    auto set_members = [&](const Keyword &k, bool nodes, const Scope &sc, std::vector<int> &members) -> int {
        const int count = nodes ? nd : ne;
        const int off = nodes ? sc.node_off : sc.elem_off;
        auto &pool = nodes ? nsets : elsets;
        const bool gen = k.params.count("generate") > 0;
        for (const auto &row : k.data) {
            if (gen) {
                long a, b, st = 1;
                if (row.size() < 2 || !to_int(row[0], a) || !to_int(row[1], b) || (row.size() > 2 && !to_int(row[2], st)) ||
                    st <= 0)
                    return fail(OFEA_ERR_IO, "bad generate line in *" + k.name);
                for (long i = a; i <= b; i += st) {
                    if (i < 1 || i + off > count) return fail(OFEA_ERR_IO, "generated id out of range in *" + k.name);
                    members.push_back((int)(i + off) - 1);
                }
                continue;
            }
            for (const auto &tok : row) {
                if (tok.empty()) continue;
                std::vector<int> r;
                int rc2 = resolve(pool, tok, count, r, nodes ? "node" : "element", sc, &inst, &parts, nodes);
                if (rc2) return rc2;
                members.insert(members.end(), r.begin(), r.end());
            }
        }
        return OFEA_OK;
    };
    // node set or node id of a *Boundary / *Cload / *Initial Conditions line
    auto resolve_nodes = [&](const std::string &tok, std::vector<int> &out) {
        return resolve(nsets, tok, nd, out, "node", scope, &inst, &parts, true);
    };
    for (const auto &k : kws) {
        if (k.name == "part") {
            auto nm = k.params.find("name");
            scope.part = nm == k.params.end() ? std::string("part") : lower(nm->second);
            scope.node_off = node_rows;
            scope.elem_off = elem_rows;
            parts[scope.part] = PartInfo{node_rows, elem_rows};
        } else if (k.name == "end part") {
            scope = Scope();
        } else if (k.name == "instance") {
            auto nm = k.params.find("name");
            auto pt = k.params.find("part");
            if (nm != k.params.end() && pt != k.params.end()) inst[lower(nm->second)] = lower(pt->second);
        } else if (k.name == "node") {
            node_rows += (int)k.data.size();
        } else if (k.name == "element") {
            elem_rows += (int)k.data.size();
        } else if (k.name == "nset" || k.name == "elset") {
            const bool nodes = k.name == "nset";
            auto it = k.params.find(nodes ? "nset" : "elset");
            if (it == k.params.end()) continue;
            // an assembly-level set of one instance ("instance=I") takes labels of that part
            Scope sc = scope;
            auto in = k.params.find("instance");
            if (scope.part.empty() && in != k.params.end()) {
                auto ii = inst.find(lower(in->second));
                if (ii == inst.end()) return fail(OFEA_ERR_IO, "*" + k.name + " names an unknown instance " + in->second);
                const PartInfo &pi = parts[ii->second];
                sc.node_off = pi.node_off;
                sc.elem_off = pi.elem_off;
            }
            std::vector<int> m;
            if ((rc = set_members(k, nodes, sc, m))) return rc;
            auto &pool = nodes ? nsets : elsets;
            const std::string name = lower(it->second);
            if (!scope.part.empty()) {
                auto &q = pool[scope.part + "." + name];
                q.insert(q.end(), m.begin(), m.end());
            }
            auto &dst = pool[name];
            dst.insert(dst.end(), m.begin(), m.end());
        } else if (k.name == "surface") {
            auto ty = k.params.find("type");
            if (ty != k.params.end() && lower(ty->second) != "element") continue;
            auto nm = k.params.find("name");
            if (nm == k.params.end()) continue;
            auto &surf = surfaces[lower(nm->second)];
            for (const auto &row : k.data) {
                if (row.size() < 2) return fail(OFEA_ERR_IO, "bad *Surface line");
                std::string f = lower(row[1]);
                if (f.size() < 2 || f[0] != 's') return fail(OFEA_ERR_IO, "bad face label " + row[1]);
                const int face = std::atoi(f.c_str() + 1) - 1;
                std::vector<int> els;
                if ((rc = resolve(elsets, row[0], ne, els, "element", scope, &inst, &parts, false))) return rc;
                for (int e : els) {
                    const int nf = D.etype[e] == OFEA_C3D8 ? 6 : 4;
                    if (face < 0 || face >= nf) return fail(OFEA_ERR_IO, "face label out of range: " + row[1]);
                    surf.emplace_back(e, face);
                }
            }
        } else if (k.name == "material") {
            auto nm = k.params.find("name");
            cur_mat = nm == k.params.end() ? std::string() : lower(nm->second);
            mats[cur_mat];
        } else if (k.name == "elastic" || k.name == "density") {
            if (k.data.empty()) return fail(OFEA_ERR_IO, "*" + k.name + " without a data line");
            Mat &m = mats[cur_mat];
            double a, b;
            if (k.name == "elastic") {
                if (k.data[0].size() < 2 || !to_double(k.data[0][0], a) || !to_double(k.data[0][1], b))
                    return fail(OFEA_ERR_IO, "bad *Elastic line");
                m.E = a;
                m.nu = b;
            } else {
                if (k.data[0].empty() || !to_double(k.data[0][0], a)) return fail(OFEA_ERR_IO, "bad *Density line");
                m.rho = a;
            }
        } else if (k.name == "solid section") {
            auto es = k.params.find("elset");
            auto mt = k.params.find("material");
            if (es == k.params.end() || mt == k.params.end()) continue;
            // a section inside a part names that part's element set
            const std::string set = lower(es->second);
            const bool own = !scope.part.empty() && elsets.count(scope.part + "." + set);
            D.sections.emplace_back(own ? scope.part + "." + set : set, lower(mt->second));
        } else if (k.name == "cload" || (k.name == "initial conditions" && k.params.count("type") &&
                                          lower(k.params.at("type")) == "velocity")) {
            const bool vel = k.name != "cload";
            for (const auto &row : k.data) {
                long dof;
                double val;
                if (row.size() < 3 || !to_int(row[1], dof) || !to_double(row[2], val))
                    return fail(OFEA_ERR_IO, "bad *" + k.name + " line");
                if ((rc = resolve_nodes(row[0], tgt))) return rc;
                if (dof < 1 || dof > 3) continue;  // rotational DOFs do not exist on solid nodes
                for (int i : tgt) {
                    if (vel)
                        D.v0[3 * (size_t)i + dof - 1] = val;
                    else
                        D.f_nodal[3 * (size_t)i + dof - 1] += val;  // loads add
                }
            }
        } else if (k.name == "boundary") {
            for (const auto &row : k.data) {
                if (row.size() < 2) return fail(OFEA_ERR_IO, "bad *Boundary line");
                if ((rc = resolve_nodes(row[0], tgt))) return rc;
                int mask = 0;
                long a, b;
                const std::string t = lower(row[1]);
                if (t == "encastre" || t == "pinned")
                    mask = 7;
                else if (t == "xsymm")
                    mask = 1;
                else if (t == "ysymm")
                    mask = 2;
                else if (t == "zsymm")
                    mask = 4;
                else if (to_int(row[1], a)) {
                    b = a;
                    if (row.size() > 2 && !row[2].empty() && !to_int(row[2], b)) return fail(OFEA_ERR_IO, "bad *Boundary line");
                    double v = 0.0;
                    if (row.size() > 3 && !row[3].empty() && !to_double(row[3], v))
                        return fail(OFEA_ERR_IO, "bad *Boundary line");
                    for (long d = std::max(a, 1L); d <= std::min(b, 3L); ++d) mask |= 1 << (d - 1);
                    if (v != 0.0) D.bc_nonzero += (int)tgt.size();
                } else {
                    return fail(OFEA_ERR_IO, "unsupported *Boundary type " + row[1]);
                }
                for (int i : tgt) D.bc_mask[i] |= (unsigned char)mask;
            }
        } else if (k.name == "dsload") {
            for (const auto &row : k.data) {
                if (row.size() < 3) return fail(OFEA_ERR_IO, "bad *Dsload line");
                const std::string lab = lower(row[1]);
                double mag, dir[3] = {0, 0, 0};
                if (!to_double(row[2], mag)) return fail(OFEA_ERR_IO, "bad *Dsload magnitude");
                const bool trvec = lab == "trvec";
                if (trvec) {
                    if (row.size() < 6 || !to_double(row[3], dir[0]) || !to_double(row[4], dir[1]) ||
                        !to_double(row[5], dir[2]))
                        return fail(OFEA_ERR_IO, "bad TRVEC direction");
                    const double l = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
                    if (!(l > 0.0)) return fail(OFEA_ERR_IO, "zero TRVEC direction");
                    for (double &x : dir) x /= l;
                } else if (lab != "p") {
                    return fail(OFEA_ERR_IO, "unsupported *Dsload label " + row[1]);
                }
                auto it = surfaces.find(lower(row[0]));
                if (it == surfaces.end()) return fail(OFEA_ERR_IO, "unknown surface " + row[0]);
                for (auto [e, face] : it->second) {
                    const bool hex = D.etype[e] == OFEA_C3D8;
                    const int nn = hex ? 8 : 4, nf = hex ? 4 : 3;
                    const int *fl = hex ? HEX_FACES[face] : TET_FACES[face];
                    const int *cn = D.conn.data() + 8 * (size_t)e;
                    const double *P[4];
                    for (int a = 0; a < nf; ++a) P[a] = D.xyz.data() + 3 * (size_t)(cn[fl[a]] - 1);
                    // centroid of the element: the outward normal points away from it
                    double cx[3] = {0, 0, 0};
                    for (int a = 0; a < nn; ++a)
                        for (int c = 0; c < 3; ++c) cx[c] += D.xyz[3 * (size_t)(cn[a] - 1) + c] / nn;
                    // per face node: integral of N_a dA times the area-weighted normal
                    // (w: integral of N_a times the area-weighted normal; wa: integral of N_a dA)
                    double w[4][3], wa[4];
                    if (!hex) {
                        Cross n = cross(sub(P[1], P[0]), sub(P[2], P[0]));  // 2 A n
                        const double a2 = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
                        for (int a = 0; a < 3; ++a) {
                            w[a][0] = n.x / 6.0;
                            w[a][1] = n.y / 6.0;
                            w[a][2] = n.z / 6.0;
                            wa[a] = a2 / 6.0;
                        }
                    } else {  // bilinear quad, 2x2 Gauss: sum_g N_a(g) (x_xi x x_eta)(g)
                        const double q = 1.0 / std::sqrt(3.0);
                        const double xs[4] = {-1, 1, 1, -1}, es[4] = {-1, -1, 1, 1};
                        for (int a = 0; a < 4; ++a) w[a][0] = w[a][1] = w[a][2] = wa[a] = 0.0;
                        for (int g = 0; g < 4; ++g) {
                            const double xi = xs[g] * q, eta = es[g] * q;
                            double dxi[3] = {0, 0, 0}, deta[3] = {0, 0, 0}, N[4];
                            for (int a = 0; a < 4; ++a) {
                                N[a] = 0.25 * (1 + xs[a] * xi) * (1 + es[a] * eta);
                                for (int c = 0; c < 3; ++c) {
                                    dxi[c] += 0.25 * xs[a] * (1 + es[a] * eta) * P[a][c];
                                    deta[c] += 0.25 * es[a] * (1 + xs[a] * xi) * P[a][c];
                                }
                            }
                            Cross n = cross({dxi[0], dxi[1], dxi[2]}, {deta[0], deta[1], deta[2]});
                            const double dA = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
                            for (int a = 0; a < 4; ++a) {
                                w[a][0] += N[a] * n.x;
                                w[a][1] += N[a] * n.y;
                                w[a][2] += N[a] * n.z;
                                wa[a] += N[a] * dA;
                            }
                        }
                    }
                    // orient: total area vector away from the element centroid
                    double tot[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
                    for (int a = 0; a < nf; ++a)
                        for (int c = 0; c < 3; ++c) {
                            tot[c] += w[a][c];
                            fc[c] += P[a][c] / nf;
                        }
                    const double s = (tot[0] * (fc[0] - cx[0]) + tot[1] * (fc[1] - cx[1]) + tot[2] * (fc[2] - cx[2])) < 0
                                         ? -1.0 : 1.0;
                    for (int a = 0; a < nf; ++a) {  // pressure: t = -p n_out; TRVEC: t = mag d
                        const int node = cn[fl[a]] - 1;
                        for (int c = 0; c < 3; ++c)
                            D.f_nodal[3 * (size_t)node + c] += trvec ? mag * dir[c] * wa[a] : -mag * s * w[a][c];
                    }
                }
            }
        }
    }
    for (const auto &sec : D.sections) {
        auto es = elsets.find(sec.first);
        auto mt = mats.find(sec.second);
        if (es == elsets.end()) return fail(OFEA_ERR_IO, "*Solid Section on unknown elset " + sec.first);
        if (mt == mats.end()) return fail(OFEA_ERR_IO, "*Solid Section with unknown material " + sec.second);
        for (int e : es->second) {
            D.E_elem[e] = mt->second.E;
            D.nu_elem[e] = mt->second.nu;
            D.rho_elem[e] = mt->second.rho;
        }
    }
    D.ext_ok = true;
    return OFEA_OK;
}

}  // namespace ofea

using namespace ofea;

extern "C" {

static int ext_ready(const ofea_deck *D)
{
    if (!D) return fail(OFEA_ERR_ARG, "null deck");
    if (!D->ext_ok) return fail(OFEA_ERR_IO, "deck extensions unreadable: " + D->ext_error);
    return OFEA_OK;
}

int ofea_inp_nodal_loads(const ofea_deck *D, double *f)
{
    int rc = ext_ready(D);
    if (rc) return rc;
    if (!f) return fail(OFEA_ERR_ARG, "null output");
    std::memcpy(f, D->f_nodal.data(), D->f_nodal.size() * sizeof(double));
    return OFEA_OK;
}

int ofea_inp_boundary(const ofea_deck *D, int *nfixed, int *fixed, int *npartial, int *nnonzero)
{
    int rc = ext_ready(D);
    if (rc) return rc;
    int nf = 0, np = 0;
    for (int i = 0; i < D->nd; ++i) {
        const int m = D->bc_mask[i];
        if (m == 7) {
            if (fixed) fixed[nf] = i + 1;
            ++nf;
        } else {
            np += (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1);
        }
    }
    if (nfixed) *nfixed = nf;
    if (npartial) *npartial = np;
    if (nnonzero) *nnonzero = D->bc_nonzero;
    return OFEA_OK;
}

int ofea_inp_initial_velocity(const ofea_deck *D, double *v0)
{
    int rc = ext_ready(D);
    if (rc) return rc;
    if (!v0) return fail(OFEA_ERR_ARG, "null output");
    std::memcpy(v0, D->v0.data(), D->v0.size() * sizeof(double));
    return OFEA_OK;
}

int ofea_inp_materials(const ofea_deck *D, double *E_elem, double *nu_elem, double *rho_elem)
{
    int rc = ext_ready(D);
    if (rc) return rc;
    if (E_elem) std::memcpy(E_elem, D->E_elem.data(), D->E_elem.size() * sizeof(double));
    if (nu_elem) std::memcpy(nu_elem, D->nu_elem.data(), D->nu_elem.size() * sizeof(double));
    if (rho_elem) std::memcpy(rho_elem, D->rho_elem.data(), D->rho_elem.size() * sizeof(double));
    return OFEA_OK;
}

}  // extern "C"
