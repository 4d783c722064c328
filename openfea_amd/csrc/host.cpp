// host.cpp -- host-side drop-in pieces of libofea_hip.so: the reference-dialect
// .inp reader, the boundary-condition numbering, the CSC pattern builder and
// the column-owner gather plans that the assembly kernels consume.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>

#include "ofea_internal.h"

namespace ofea {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// ---------------------------------------------------------------- .inp reader
// Line reader accepting \n, \r\n and \r endings (ReadInfo::safeGetline,
// src/data2cae.cpp:821-851).  Returns false at end of file with no characters.
static bool get_line(std::istream &is, std::string &t)
{
    t.clear();
    std::streambuf *sb = is.rdbuf();
    for (;;) {
        int c = sb->sbumpc();
        if (c == '\n') return true;
        if (c == '\r') {
            if (sb->sgetc() == '\n') sb->sbumpc();
            return true;
        }
        if (c == std::streambuf::traits_type::eof()) {
            is.setstate(std::ios::eofbit);
            return !t.empty();
        }
        t += (char)c;
    }
}

static std::string strip_blanks(std::string s)
{
    s.erase(std::remove(s.begin(), s.end(), ' '), s.end());
    return s;
}

static void drop_last(std::string &s)
{
    if (!s.empty()) s.erase(s.end() - 1);
}

// ReadInfo::read_ele_node_num (src/data2cae.cpp:18-53): the second token of the
// line after the first "*Nset" / "*Elset" keyword lines.
static int read_counts(const std::string &path, int &nd, int &ne)
{
    std::ifstream in(path);
    if (!in) return fail(OFEA_ERR_IO, "cannot open " + path);
    std::string line;
    nd = ne = -1;
    while (get_line(in, line)) {
        if (line.find("*Nset") != std::string::npos) {
            get_line(in, line);
            std::istringstream iss(line);
            std::string a, b, c;
            iss >> a >> b >> c;
            nd = std::atoi(b.c_str());
        }
        if (line.find("*Elset") != std::string::npos) {
            get_line(in, line);
            std::istringstream iss(line);
            std::string a, b, c;
            iss >> a >> b >> c;
            ne = std::atoi(b.c_str());
            break;
        }
    }
    if (nd <= 0 || ne <= 0) return fail(OFEA_ERR_IO, "missing *Nset/*Elset count lines in " + path);
    return OFEA_OK;
}

// ReadInfo::read_geo_mesh (src/data2cae.cpp:77-179): node and element order is
// line order; "id, x, y, z" with the trailing commas of x and y dropped;
// element lines "id, n1, ..., nk"; the element type comes from "type=" and an
// unknown name maps to C3D4 (the std::map operator[] default of ELE_TYPES).
static int read_mesh(const std::string &path, ofea_deck &D)
{
    std::ifstream in(path);
    if (!in) return fail(OFEA_ERR_IO, "cannot open " + path);
    D.xyz.assign((size_t)D.nd * 3, 0.0);
    D.conn.assign((size_t)D.ne * 8, 0);
    D.etype.assign((size_t)D.ne, 0);
    std::map<std::string, int> types = {{"C3D4", OFEA_C3D4}, {"C3D8", OFEA_C3D8}};
    std::string line;
    bool node_rec = false, elem_rec = false;
    int id_node = 0, id_ele = 0, cur_type = 0;
    while (get_line(in, line)) {
        if (line.find("*Node") != std::string::npos) {
            node_rec = true;
            get_line(in, line);
        }
        if (node_rec && line.find('*') != std::string::npos) node_rec = false;
        if (node_rec) {
            std::istringstream iss(line);
            std::string id, x, y, z;
            iss >> id >> x >> y >> z;
            drop_last(x);
            drop_last(y);
            if (id_node >= D.nd) return fail(OFEA_ERR_IO, "more node lines than the *Nset count");
            try {
                D.xyz[3 * id_node + 0] = std::stod(x);
                D.xyz[3 * id_node + 1] = std::stod(y);
                D.xyz[3 * id_node + 2] = std::stod(z);
            } catch (...) {
                return fail(OFEA_ERR_IO, "bad node line: " + line);
            }
            id_node++;
        }
        if (line.find("*Element") != std::string::npos) {
            size_t eq = line.find('=');
            std::string name = eq == std::string::npos ? std::string() : strip_blanks(line.substr(eq + 1));
            cur_type = types[name];
            elem_rec = true;
            get_line(in, line);
        }
        if (elem_rec && line.find('*') != std::string::npos) elem_rec = false;
        if (elem_rec) {
            if (id_ele >= D.ne) return fail(OFEA_ERR_IO, "more element lines than the *Elset count");
            D.etype[id_ele] = cur_type;
            std::istringstream iss(line);
            std::string tok;
            int slot = -1;
            while (std::getline(iss, tok, ',')) {
                tok = strip_blanks(tok);
                if (slot >= 0 && slot < 8) D.conn[(size_t)8 * id_ele + slot] = std::atoi(tok.c_str());
                slot++;
            }
            id_ele++;
        }
        if (line.find("*End Assembly") != std::string::npos) break;
    }
    if (id_node != D.nd || id_ele != D.ne) return fail(OFEA_ERR_IO, "node/element line count differs from *Nset/*Elset");
    return OFEA_OK;
}

// Node-set scan of ReadInfo::read_load_bcs / read_dis_bcs
// (src/data2cae.cpp:315-414): ids on the lines after the set keyword line, up
// to the next line holding '*'; the scan stops at "*End Assembly".
static void read_set(std::istream &in, const std::string &key, std::vector<int> &out)
{
    std::string line;
    while (get_line(in, line)) {
        if (line.find(key) != std::string::npos) {
            while (get_line(in, line)) {
                if (line.find('*') != std::string::npos) break;
                std::istringstream iss(line);
                std::string tok;
                while (iss >> tok) {
                    if (tok.find(',') != std::string::npos) drop_last(tok);
                    out.push_back(std::atoi(tok.c_str()));
                }
            }
        }
        if (line.find("*End Assembly") != std::string::npos) break;
    }
}

static int read_bcs_mat_time(const std::string &path, ofea_deck &D)
{
    {
        std::ifstream in(path);
        if (!in) return fail(OFEA_ERR_IO, "cannot open " + path);
        read_set(in, "Set-load", D.load_set);
        // first "Cload" keyword after the sets: "set, dof, value" (src/data2cae.cpp:352-370)
        std::string line;
        while (get_line(in, line)) {
            if (line.find("Cload") != std::string::npos) {
                get_line(in, line);
                std::istringstream iss(line);
                std::string name, dof, val;
                iss >> name >> dof >> val;
                drop_last(dof);
                D.load_dof = std::atoi(dof.c_str());
                try {
                    D.load_value = std::stod(val);
                } catch (...) {
                    return fail(OFEA_ERR_IO, "bad *Cload line: " + line);
                }
                break;
            }
        }
    }
    {
        std::ifstream in(path);
        read_set(in, "Set-fix", D.fix_set);
    }
    {
        // ReadInfo::read_mat (src/data2cae.cpp:416-446)
        std::ifstream in(path);
        std::string line;
        while (get_line(in, line)) {
            if (line.find("*Density") != std::string::npos) {
                get_line(in, line);
                std::istringstream iss(line);
                std::string rho;
                iss >> rho;
                try { D.rho = std::stod(rho); } catch (...) { return fail(OFEA_ERR_IO, "bad *Density line"); }
            }
            if (line.find("*Elastic") != std::string::npos) {
                get_line(in, line);
                std::istringstream iss(line);
                std::string e, nu;
                iss >> e >> nu;
                drop_last(e);
                try {
                    D.E = std::stod(e);
                    D.nu = std::stod(nu);
                } catch (...) {
                    return fail(OFEA_ERR_IO, "bad *Elastic line");
                }
            }
        }
    }
    {
        // ReadInfo::read_time (src/data2cae.cpp:448-492): "dt, T" after the
        // "Explicit" keyword line; an omitted dt leaves 0 (automatic step).
        std::ifstream in(path);
        std::string line;
        while (get_line(in, line)) {
            if (line.find("Explicit") == std::string::npos) continue;
            while (get_line(in, line)) {
                if (line.size() >= 2 && line[0] == '*' && line[1] == '*') continue;
                if (!line.empty() && line[0] == '*') break;
                std::istringstream iss(line);
                double dt;
                if (iss >> dt)
                    D.dt = dt;
                else {
                    iss.clear();
                    iss.seekg(0);
                }
                char comma;
                iss >> comma;
                double T;
                if (iss >> T) D.t_total = T;
            }
        }
    }
    return OFEA_OK;
}

// ---------------------------------------------------------------- pattern
int build_pattern(int nd, int ne, const int *conn, const int *etype, const int *resort, Pattern &P)
{
    int nfree = 0;
    for (int i = 0; i < nd; ++i)
        if (resort[i] >= 0) nfree = std::max(nfree, resort[i] + 1);
    std::vector<int> rank_node(nfree, -1);
    for (int i = 0; i < nd; ++i)
        if (resort[i] >= 0) rank_node[resort[i]] = i;
    // node -> element incidence
    std::vector<int> inc_ptr(nd + 1, 0);
    for (int e = 0; e < ne; ++e) {
        int nn = etype[e] == OFEA_C3D8 ? 8 : 4;
        for (int a = 0; a < nn; ++a) {
            int node = conn[8 * e + a] - 1;
            if (node < 0 || node >= nd) return fail(OFEA_ERR_ARG, "element node id out of range");
            inc_ptr[node + 1]++;
        }
    }
    for (int i = 0; i < nd; ++i) inc_ptr[i + 1] += inc_ptr[i];
    std::vector<int> inc(inc_ptr[nd]);
    {
        std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int e = 0; e < ne; ++e) {
            int nn = etype[e] == OFEA_C3D8 ? 8 : 4;
            for (int a = 0; a < nn; ++a) inc[fill[conn[8 * e + a] - 1]++] = e;
        }
    }
    P.nfree = nfree;
    P.n = 3 * nfree;
    P.node_blk.assign(nfree + 1, 0);
    std::vector<int> stamp(nfree, -1), nb;
    std::vector<std::vector<int>> adj(nfree);
    for (int r = 0; r < nfree; ++r) {
        int node = rank_node[r];
        nb.clear();
        if (node >= 0)
            for (int t = inc_ptr[node]; t < inc_ptr[node + 1]; ++t) {
                int e = inc[t];
                int nn = etype[e] == OFEA_C3D8 ? 8 : 4;
                for (int b = 0; b < nn; ++b) {
                    int s = resort[conn[8 * e + b] - 1];
                    if (s >= 0 && stamp[s] != r) {
                        stamp[s] = r;
                        nb.push_back(s);
                    }
                }
            }
        std::sort(nb.begin(), nb.end());
        adj[r] = nb;
        P.node_blk[r + 1] = P.node_blk[r] + (int)nb.size();
    }
    long long nblk = P.node_blk[nfree];
    if (9 * nblk > (long long)std::numeric_limits<int>::max())
        return fail(OFEA_ERR_LIMIT, "nnz exceeds the int32 index range of the CSC arrays");
    P.nnz = 9 * nblk;
    P.col_ptr.assign(P.n + 1, 0);
    P.row_idx.resize(P.nnz);
    P.bcol.resize(nblk);
    P.diag_k.assign(nfree, -1);
    for (int r = 0; r < nfree; ++r) {
        int deg = (int)adj[r].size();
        auto it = std::lower_bound(adj[r].begin(), adj[r].end(), r);
        if (it != adj[r].end() && *it == r) P.diag_k[r] = (int)(it - adj[r].begin());
        int base = 9 * P.node_blk[r];
        for (int k = 0; k < deg; ++k) P.bcol[P.node_blk[r] + k] = adj[r][k];
        for (int c = 0; c < 3; ++c) {
            int cb = base + c * 3 * deg;
            P.col_ptr[3 * r + c] = cb;
            for (int k = 0; k < deg; ++k)
                for (int d = 0; d < 3; ++d) P.row_idx[cb + 3 * k + d] = 3 * adj[r][k] + d;
        }
    }
    P.col_ptr[P.n] = (int)P.nnz;
    return OFEA_OK;
}

int build_gather_plan(int nd, int ne, const int *conn, const int *etype, const int *resort, const Pattern &P,
                      int type, int pair_cap, GatherPlan &G, bool sparse)
{
    G = GatherPlan();
    G.nn = type == OFEA_C3D8 ? 8 : 4;
    const int nn = G.nn;
    std::vector<int> local(ne, -1);
    for (int e = 0; e < ne; ++e)
        if (etype[e] == type) {
            local[e] = G.ne++;
            G.elem_global.push_back(e);
            for (int a = 0; a < nn; ++a) G.conn.push_back(conn[8 * e + a] - 1);
        }
    const int nfree = P.nfree;
    // per free node: incident elements of this type (ascending) with local index
    std::vector<int> cnt(nfree + 1, 0);
    for (int e = 0; e < ne; ++e) {
        if (etype[e] != type) continue;
        for (int a = 0; a < nn; ++a) {
            int r = resort[conn[8 * e + a] - 1];
            if (r >= 0) cnt[r + 1]++;
        }
    }
    for (int r = 0; r < nfree; ++r) cnt[r + 1] += cnt[r];
    std::vector<int> pairs(cnt[nfree]);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int e = 0; e < ne; ++e) {  // ascending e => ascending per node
            if (etype[e] != type) continue;
            for (int a = 0; a < nn; ++a) {
                int r = resort[conn[8 * e + a] - 1];
                if (r >= 0) pairs[fill[r]++] = (local[e] << 3) | a;
            }
        }
    }
    if ((long long)G.ne >= (1LL << 28)) return fail(OFEA_ERR_LIMIT, "too many elements for the packed pair index");
    if (nfree >= (1 << 23)) return fail(OFEA_ERR_LIMIT, "too many free nodes for the packed block index");
    G.pair = pairs;
    G.node_pair = cnt;
    G.chunk_node.push_back(0);
    G.chunk_pair.push_back(0);
    G.chunk_blk.push_back(0);
    G.contrib_ptr.push_back(0);
    G.off_ptr.push_back(0);
    int chunk_first_pair = 0, chunk_first_node = 0;
    int in_chunk = 0;
    std::vector<std::vector<uint16_t>> lists;
    for (int r = 0; r < nfree; ++r) {
        int np = cnt[r + 1] - cnt[r];
        if (np > pair_cap) return fail(OFEA_ERR_LIMIT, "a node has more incident elements than the gather chunk holds");
        int deg = P.node_blk[r + 1] - P.node_blk[r];
        if (deg > 255) return fail(OFEA_ERR_LIMIT, "a node has more than 255 neighbours");
        if (in_chunk + np > pair_cap || r - chunk_first_node >= PLAN_CHUNK_NODES) {
            G.chunk_pair.push_back(cnt[r]);
            G.chunk_blk.push_back((int)G.blk_rk.size());
            G.chunk_node.push_back(r);
            chunk_first_pair = cnt[r];
            chunk_first_node = r;
            in_chunk = 0;
        }
        in_chunk += np;
        if (sparse && np == 0) continue;  // a node no listed element touches: no blocks
        const int *nbr = P.bcol.data() + P.node_blk[r];
        lists.assign(deg, std::vector<uint16_t>());
        for (int t = cnt[r]; t < cnt[r + 1]; ++t) {
            int el = pairs[t] >> 3, a = pairs[t] & 7;
            for (int b = 0; b < nn; ++b) {
                int s = resort[G.conn[(size_t)el * nn + b]];
                if (s < 0) continue;
                int k = (int)(std::lower_bound(nbr, nbr + deg, s) - nbr);
                if (k == deg || nbr[k] != s) return fail(OFEA_ERR_ARG, "an element couples nodes outside the pattern");
                // C3D4 kernels keep the pair's gradients rotated so that slot 0 is the owned node
                int slot = nn == 4 ? ((b - a) & 3) : b;
                lists[k].push_back((uint16_t)((t - chunk_first_pair) * nn + slot));
            }
        }
        for (int k = 0; k < deg; ++k) {
            if (sparse && lists[k].empty()) continue;  // a block no element of this list touches
            G.blk_rk.push_back((r << 8) | k);
            G.blk_rl.push_back((uint8_t)(r - chunk_first_node));
            if (sparse) G.blk_gid.push_back(P.node_blk[r] + k);
            G.contrib.insert(G.contrib.end(), lists[k].begin(), lists[k].end());
            G.contrib_ptr.push_back((int)G.contrib.size());
            if (nbr[k] != r) G.off_contrib.insert(G.off_contrib.end(), lists[k].begin(), lists[k].end());
            G.off_ptr.push_back((int)G.off_contrib.size());
        }
    }
    G.chunk_pair.push_back(cnt[nfree]);
    G.chunk_blk.push_back((int)G.blk_rk.size());
    G.chunk_node.push_back(nfree);
    (void)nd;
    return OFEA_OK;
}

int build_tet_chunks(int nd, const Pattern &P, const GatherPlan &G, const int *resort, TetChunkPlan &T)
{
    T = TetChunkPlan();
    const int nf = P.nfree;
    std::vector<int> eslot(G.ne, -1);  // chunk element slot while the chunk is built
    std::vector<int> nslot(nd, -1);    // node-table slot while the chunk is built
    // a block with c contributors takes the smallest power of two >= ceil(c / TC_CPL) lanes
    auto group_of = [](int c) {
        const int need = (c + TC_CPL - 1) / TC_CPL;
        int g = 1;
        while (g < need) g <<= 1;
        return g;
    };
    struct Block {
        int q, k, g;
        std::vector<uint32_t> ent;  // type-local element << 4 | slot_col << 2 | slot_row, ascending element
    };
    // the blocks of node q (slot k order), contributors in ascending element order
    auto node_blocks = [&](int q, std::vector<Block> &out) {
        const int p0 = G.node_pair[q], p1 = G.node_pair[q + 1];
        const int nb = P.node_blk[q], deg = P.node_blk[q + 1] - nb;
        const int *nbr = P.bcol.data() + nb;
        out.assign(deg, Block());
        for (int k = 0; k < deg; ++k) out[k].q = q, out[k].k = k;
        for (int p = p0; p < p1; ++p) {  // pairs ascend by element
            const int el = G.pair[p] >> 3, a = G.pair[p] & 7;
            for (int b = 0; b < 4; ++b) {
                const int s = resort[G.conn[4 * (size_t)el + b]];
                if (s < 0) continue;
                const int k = (int)(std::lower_bound(nbr, nbr + deg, s) - nbr);
                out[k].ent.push_back((uint32_t)el << 4 | (uint32_t)a << 2 | (uint32_t)b);
            }
        }
        for (auto &B : out) B.g = group_of((int)B.ent.size());
    };
    std::vector<Block> blocks, nodeb;
    std::vector<int> els, nodes;
    int r = 0;
    while (r < nf) {
        // grow the chunk node by node while its elements, nodes and lanes fit
        blocks.clear();
        els.clear();
        nodes.clear();
        int lanes = 0;
        const int r0 = r;
        while (r < nf) {
            node_blocks(r, nodeb);
            int lanes_r = 0, gmax = 1;
            for (const auto &B : nodeb) lanes_r += B.g, gmax = std::max(gmax, B.g);
            int add_e = 0, add_n = 0;
            for (int p = G.node_pair[r]; p < G.node_pair[r + 1]; ++p) {
                const int el = G.pair[p] >> 3;
                if (eslot[el] >= 0) continue;
                eslot[el] = -2;  // counted
                ++add_e;
                for (int b = 0; b < 4; ++b) {
                    const int v = G.conn[4 * (size_t)el + b];
                    if (nslot[v] == -1) nslot[v] = -2, ++add_n;
                }
            }
            const bool fits = gmax <= TC_GMAX && (int)els.size() + add_e <= TC_ECAP &&
                              (int)nodes.size() + add_n <= TC_NTAB && lanes + lanes_r <= TC_WG;
            // take the counted marks back, then record what the node adds if it fits
            for (int p = G.node_pair[r]; p < G.node_pair[r + 1]; ++p) {
                const int el = G.pair[p] >> 3;
                if (eslot[el] != -2) continue;
                eslot[el] = fits ? 0 : -1;
                if (fits) els.push_back(el);
                for (int b = 0; b < 4; ++b) {
                    const int v = G.conn[4 * (size_t)el + b];
                    if (nslot[v] != -2) continue;
                    nslot[v] = fits ? 0 : -1;
                    if (fits) nodes.push_back(v);
                }
            }
            if (!fits) {
                if (r == r0) {
                    for (int el : els) eslot[el] = -1;
                    for (int v : nodes) nslot[v] = -1;
                    T = TetChunkPlan();
                    return OFEA_ERR_LIMIT;  // no message: the caller falls back, set_mesh succeeds
                }
                break;
            }
            for (auto &B : nodeb) blocks.push_back(std::move(B));
            lanes += lanes_r;
            ++r;
        }
        std::sort(els.begin(), els.end());
        std::sort(nodes.begin(), nodes.end());
        for (int i = 0; i < (int)els.size(); ++i) eslot[els[i]] = i;
        for (int i = 0; i < (int)nodes.size(); ++i) nslot[nodes[i]] = i;
        const size_t n_off = T.ntab.size(), e_off = T.eloc.size(), t_off = T.task.size();
        T.ntab.resize(n_off + TC_NTAB, -1);
        T.eloc.resize(e_off + TC_ECAP, ~0u);
        T.eid.resize(e_off + TC_ECAP, -1);
        for (int i = 0; i < (int)nodes.size(); ++i) T.ntab[n_off + i] = nodes[i];
        for (int i = 0; i < (int)els.size(); ++i) {
            uint32_t w = 0;
            for (int b = 0; b < 4; ++b) w |= (uint32_t)nslot[G.conn[4 * (size_t)els[i] + b]] << (8 * b);
            T.eloc[e_off + i] = w;
            T.eid[e_off + i] = els[i];
        }
        // lanes: groups first, largest first, then single lanes by decreasing
        // contributor count, so that lane 0 of a wave has the wave's longest run
        std::stable_sort(blocks.begin(), blocks.end(), [](const Block &x, const Block &y) {
            return x.g != y.g ? x.g > y.g : (x.g == 1 && x.ent.size() > y.ent.size());
        });
        const uint32_t pad = (uint32_t)TC_ECAP;  // the zero element, slots 0 / 0
        T.task.resize(t_off + 4 * TC_WG);
        for (int l = 0; l < TC_WG; ++l) {
            T.task[t_off + 4 * l] = 0;
            for (int w = 1; w < 4; ++w) T.task[t_off + 4 * l + w] = pad | pad << 16;
        }
        const int b0 = P.node_blk[r0];
        int lane = 0;
        for (const Block &B : blocks) {
            const int c = (int)B.ent.size();
            int glog = 0;
            while ((1 << glog) < B.g) ++glog;
            const int nbl = P.node_blk[B.q] - b0, deg = P.node_blk[B.q + 1] - P.node_blk[B.q];
            for (int j = 0; j < B.g; ++j, ++lane) {
                uint32_t *rec = &T.task[t_off + 4 * (size_t)lane];
                const int i0 = (j * c) / B.g, i1 = ((j + 1) * c) / B.g;  // <= TC_CPL entries
                rec[0] = (uint32_t)B.k | (uint32_t)nbl << 8 | (uint32_t)deg << 16 | (uint32_t)glog << 24 |
                         (uint32_t)(j == 0) << 27 | (uint32_t)(i1 - i0) << 28 |
                         (uint32_t)(P.bcol[P.node_blk[B.q] + B.k] == B.q) << 31;
                for (int i = i0; i < i1; ++i) {
                    const uint32_t e = B.ent[i];
                    const uint32_t v = (uint32_t)eslot[e >> 4] | ((e >> 2) & 3u) << 8 | (e & 3u) << 10;
                    const int u = i - i0;
                    uint32_t &w = rec[1 + (u >> 1)];
                    w = (u & 1) ? ((w & 0xFFFFu) | v << 16) : ((w & 0xFFFF0000u) | v);
                }
            }
        }
        // Per step, the lanes of a quarter wave read LDS together; order each
        // lane's run so that they name different element slots modulo 16 (the
        // LDS bank of every row of a slot when the row stride is a multiple
        // of 16 doubles).  Greedy per step: each lane takes, of its remaining
        // entries, the first one whose residue is least used at that step.
        for (int q0 = 0; q0 < TC_WG; q0 += 16) {
            std::vector<uint16_t> rem[16];
            int steps = 0;
            for (int l = 0; l < 16; ++l) {
                const uint32_t *rec = &T.task[t_off + 4 * (size_t)(q0 + l)];
                const int n = (rec[0] >> 28) & 7u;
                for (int u = 0; u < n; ++u) rem[l].push_back((uint16_t)(rec[1 + (u >> 1)] >> (16 * (u & 1))));
                steps = std::max(steps, n);
            }
            std::vector<uint16_t> out[16];
            for (int st = 0; st < steps; ++st) {
                int used[16] = {0};
                for (int l = 0; l < 16; ++l)
                    if (rem[l].empty()) used[(int)pad & 15] = 1;  // the zero element: one address
                for (int l = 0; l < 16; ++l) {
                    if (rem[l].empty()) continue;
                    size_t best = 0;
                    for (size_t j = 1; j < rem[l].size(); ++j)
                        if (used[rem[l][j] & 15] < used[rem[l][best] & 15]) best = j;
                    used[rem[l][best] & 15]++;
                    out[l].push_back(rem[l][best]);
                    rem[l].erase(rem[l].begin() + (long)best);
                }
            }
            for (int l = 0; l < 16; ++l) {
                uint32_t *rec = &T.task[t_off + 4 * (size_t)(q0 + l)];
                for (int w = 1; w < 4; ++w) rec[w] = pad | pad << 16;
                for (size_t u = 0; u < out[l].size(); ++u) {
                    uint32_t &w = rec[1 + (u >> 1)];
                    w = (u & 1) ? ((w & 0xFFFFu) | (uint32_t)out[l][u] << 16) : ((w & 0xFFFF0000u) | out[l][u]);
                }
            }
        }
        // every lane of a wave carries the wave's longest run (bits 12-14 of
        // its first contributor entry), the trip count of its summation loop
        for (int w0 = 0; w0 < TC_WG; w0 += 64) {
            uint32_t trip = 0;
            for (int l = w0; l < w0 + 64; ++l) trip = std::max(trip, (T.task[t_off + 4 * (size_t)l] >> 28) & 7u);
            for (int l = w0; l < w0 + 64; ++l) T.task[t_off + 4 * (size_t)l + 1] |= trip << 12;
        }
        T.hdr.push_back(b0);
        T.hdr.push_back(P.node_blk[r] - b0);
        for (int el : els) eslot[el] = -1;
        for (int v : nodes) nslot[v] = -1;
        T.nchunks++;
    }
    return OFEA_OK;
}

// ---------------------------------------------------------------- CA change file
// ReadInfo::CA_read_change (src/data2cae.cpp:519-665): "*nodes_all, N" then
// "x, y, z" lines; "*eles_change, M, type=T" then "kind, n1, ..., nk" lines
// (kind 0 delete, 1 move, 2 add; node ids 0-based into the node list); both
// end at the next '*' line, the scan at "*End Assembly"; then "*idref2new, N"
// and one new-model node index per line (-1 = removed) up to "*End File".
static int read_ca_info(const std::string &path, ofea_ca_info &I)
{
    std::ifstream in(path);
    if (!in) return fail(OFEA_ERR_IO, "cannot open " + path);
    std::map<std::string, int> types = {{"C3D4", OFEA_C3D4}, {"C3D8", OFEA_C3D8}};
    std::string line;
    bool node_rec = false, elem_rec = false, map_rec = false;
    int cur_type = OFEA_C3D4;
    while (get_line(in, line)) {
        if (map_rec) {
            if (line.find("*End File") != std::string::npos) break;
            I.idref2new.push_back(std::atoi(line.c_str()));
            continue;
        }
        if (line.find("*idref2new") != std::string::npos) {
            map_rec = true;
            node_rec = elem_rec = false;
            continue;
        }
        if (line.find("*End Assembly") != std::string::npos) {
            node_rec = elem_rec = false;
            continue;
        }
        if (line.find("*nodes_all") != std::string::npos) {
            node_rec = true;
            elem_rec = false;
            continue;
        }
        if (line.find("*eles_change") != std::string::npos) {
            size_t eq = line.find('=');
            cur_type = eq == std::string::npos ? OFEA_C3D4 : types[strip_blanks(line.substr(eq + 1))];
            elem_rec = true;
            node_rec = false;
            continue;
        }
        if (line.find('*') != std::string::npos) {
            node_rec = elem_rec = false;
            continue;
        }
        if (node_rec) {
            std::istringstream iss(line);
            std::string x, y, z;
            iss >> x >> y >> z;
            drop_last(x);
            drop_last(y);
            try {
                I.xyz.push_back(std::stod(x));
                I.xyz.push_back(std::stod(y));
                I.xyz.push_back(std::stod(z));
            } catch (...) {
                return fail(OFEA_ERR_IO, "bad node line in " + path + ": " + line);
            }
        } else if (elem_rec) {
            std::istringstream iss(line);
            std::string tok;
            std::vector<int> v;
            while (std::getline(iss, tok, ',')) v.push_back(std::atoi(strip_blanks(tok).c_str()));
            const int nn = cur_type == OFEA_C3D8 ? 8 : 4;
            if ((int)v.size() < nn + 1) return fail(OFEA_ERR_IO, "short element line in " + path + ": " + line);
            I.kind.push_back(v[0]);
            I.etype.push_back(cur_type);
            for (int a = 0; a < 8; ++a) I.conn.push_back(a < nn ? v[1 + a] : 0);
        }
    }
    if (I.xyz.empty()) return fail(OFEA_ERR_IO, "no *nodes_all section in " + path);
    return OFEA_OK;
}

}  // namespace ofea

// ====================================================================== C ABI (host-only part)
using namespace ofea;

extern "C" {

const char *ofea_last_error(void) { return g_err.c_str(); }
const char *ofea_version(void) { return "openfea_amd 0.1 gfx950"; }

int ofea_inp_open(const char *path, ofea_deck **deck)
{
    if (!path || !deck) return fail(OFEA_ERR_ARG, "null argument");
    *deck = nullptr;
    auto *D = new ofea_deck();
    D->E = D->nu = D->rho = std::numeric_limits<double>::quiet_NaN();
    int rc = read_counts(path, D->nd, D->ne);
    if (rc == OFEA_OK) rc = read_mesh(path, *D);
    if (rc == OFEA_OK) rc = read_bcs_mat_time(path, *D);
    if (rc != OFEA_OK) {
        delete D;
        return rc;
    }
    // the extensions never make a deck the reference reads unreadable: their
    // error is reported by the ofea_inp_* getters that need them
    const std::string keep = ofea_last_error();
    if (parse_inp_ext(path, *D) != OFEA_OK) {
        D->ext_ok = false;
        D->ext_error = ofea_last_error();
    }
    set_error(keep);
    *deck = D;
    return OFEA_OK;
}

int ofea_inp_sizes(const ofea_deck *D, int *nd, int *ne, int *nfix, int *nload)
{
    if (!D) return fail(OFEA_ERR_ARG, "null deck");
    if (nd) *nd = D->nd;
    if (ne) *ne = D->ne;
    if (nfix) *nfix = (int)D->fix_set.size();
    if (nload) *nload = (int)D->load_set.size();
    return OFEA_OK;
}

int ofea_inp_get(const ofea_deck *D, double *xyz, int *conn, int *etype, int *fix_set, int *load_set, int *load_dof,
                 double *load_value, double *E, double *nu, double *rho, double *dt, double *t_total)
{
    if (!D) return fail(OFEA_ERR_ARG, "null deck");
    if (xyz) std::memcpy(xyz, D->xyz.data(), D->xyz.size() * sizeof(double));
    if (conn) std::memcpy(conn, D->conn.data(), D->conn.size() * sizeof(int));
    if (etype) std::memcpy(etype, D->etype.data(), D->etype.size() * sizeof(int));
    if (fix_set && !D->fix_set.empty()) std::memcpy(fix_set, D->fix_set.data(), D->fix_set.size() * sizeof(int));
    if (load_set && !D->load_set.empty()) std::memcpy(load_set, D->load_set.data(), D->load_set.size() * sizeof(int));
    if (load_dof) *load_dof = D->load_dof;
    if (load_value) *load_value = D->load_value;
    if (E) *E = D->E;
    if (nu) *nu = D->nu;
    if (rho) *rho = D->rho;
    if (dt) *dt = D->dt;
    if (t_total) *t_total = D->t_total;
    return OFEA_OK;
}

void ofea_inp_close(ofea_deck *deck) { delete deck; }

int ofea_build_free_index(int nd, int *fix_set, int nfix, int *resort)
{
    if (nd < 0 || nfix < 0 || !resort || (nfix > 0 && !fix_set)) return fail(OFEA_ERR_ARG, "bad argument");
    std::sort(fix_set, fix_set + nfix);
    int rank = 0, cur = 0;
    for (int i = 0; i < nd; ++i) {
        if (nfix > 0 && i == fix_set[cur] - 1) {
            resort[i] = -1;
            if (cur < nfix - 1) cur++;
        } else {
            resort[i] = rank++;
        }
    }
    return rank;
}

int ofea_build_single_load(int n, const int *resort, const int *load_set, int nload, int dof, double value, double *F)
{
    if (n < 0 || !F || (nload > 0 && (!load_set || !resort))) return fail(OFEA_ERR_ARG, "bad argument");
    std::fill(F, F + n, 0.0);
    if (dof < 1 || dof > 3) return fail(OFEA_ERR_ARG, "the DOF of the load node must be 1, 2 or 3");
    for (int i = 0; i < nload; ++i) {
        int r = resort[load_set[i] - 1];
        int idx = 3 * r + dof - 1;
        if (r < 0 || idx >= n) continue;  // a clamped load node carries no free DOF
        F[idx] = value;
    }
    return OFEA_OK;
}

int ofea_build_explicit_load(int nd, const int *load_set, int nload, int dof, double value, double *f_ext)
{
    if (nd < 0 || !f_ext || nload < 0 || (nload > 0 && !load_set)) return fail(OFEA_ERR_ARG, "bad argument");
    std::fill(f_ext, f_ext + 3 * (size_t)nd, 0.0);
    if (dof < 1 || dof > 3) return fail(OFEA_ERR_ARG, "the DOF of the load node must be 1, 2 or 3");
    if (!(value > -2147483649.0 && value < 2147483648.0)) return fail(OFEA_ERR_ARG, "load value outside the int range");
    const int force_value = (int)value;  // int, as the reference holds it: the fraction is dropped
    for (int i = 0; i < nload; ++i) {
        const int node = load_set[i] - 1;
        if (node < 0 || node >= nd) return fail(OFEA_ERR_ARG, "load node id out of range");
        f_ext[3 * (size_t)node + dof - 1] = force_value;
    }
    return OFEA_OK;
}

int ofea_build_csr(int nd, int ne, const int *conn, const int *etype, const int *resort, int *n, long long *nnz,
                   int *col_ptr, int *row_idx)
{
    if (nd <= 0 || ne <= 0 || !conn || !etype || !resort) return fail(OFEA_ERR_ARG, "bad mesh arguments");
    for (int e = 0; e < ne; ++e)
        if (etype[e] != OFEA_C3D4 && etype[e] != OFEA_C3D8) return fail(OFEA_ERR_ARG, "unknown element type id");
    Pattern P;
    int rc = build_pattern(nd, ne, conn, etype, resort, P);
    if (rc) return rc;
    if (n) *n = P.n;
    if (nnz) *nnz = P.nnz;
    if (col_ptr) std::memcpy(col_ptr, P.col_ptr.data(), P.col_ptr.size() * sizeof(int));
    if (row_idx) std::memcpy(row_idx, P.row_idx.data(), P.row_idx.size() * sizeof(int));
    return OFEA_OK;
}

int ofea_reset_displacement(int nd, const int *resort, const double *u_free, double *u_full)
{
    if (nd < 0 || !resort || !u_full || !u_free) return fail(OFEA_ERR_ARG, "bad argument");
    for (int i = 0; i < nd; ++i) {
        int r = resort[i];
        for (int c = 0; c < 3; ++c) u_full[3 * i + c] = r >= 0 ? u_free[3 * r + c] : 0.0;
    }
    return OFEA_OK;
}

int ofea_ca_info_open(const char *path, ofea_ca_info **info)
{
    if (!path || !info) return fail(OFEA_ERR_ARG, "null argument");
    *info = nullptr;
    auto *I = new ofea_ca_info();
    int rc = read_ca_info(path, *I);
    if (rc != OFEA_OK) {
        delete I;
        return rc;
    }
    *info = I;
    return OFEA_OK;
}

int ofea_ca_info_sizes(const ofea_ca_info *I, int *nd_all, int *nmod, int *nmap)
{
    if (!I) return fail(OFEA_ERR_ARG, "null info");
    if (nd_all) *nd_all = (int)(I->xyz.size() / 3);
    if (nmod) *nmod = (int)I->kind.size();
    if (nmap) *nmap = (int)I->idref2new.size();
    return OFEA_OK;
}

int ofea_ca_info_get(const ofea_ca_info *I, double *xyz_all, int *kind, int *conn_mod, int *etype_mod,
                     int *idref2new)
{
    if (!I) return fail(OFEA_ERR_ARG, "null info");
    if (xyz_all) std::memcpy(xyz_all, I->xyz.data(), I->xyz.size() * sizeof(double));
    if (kind && !I->kind.empty()) std::memcpy(kind, I->kind.data(), I->kind.size() * sizeof(int));
    if (conn_mod && !I->conn.empty()) std::memcpy(conn_mod, I->conn.data(), I->conn.size() * sizeof(int));
    if (etype_mod && !I->etype.empty()) std::memcpy(etype_mod, I->etype.data(), I->etype.size() * sizeof(int));
    if (idref2new && !I->idref2new.empty())
        std::memcpy(idref2new, I->idref2new.data(), I->idref2new.size() * sizeof(int));
    return OFEA_OK;
}

void ofea_ca_info_close(ofea_ca_info *info) { delete info; }

int ofea_reset_ca_displacement(int nd_all, const int *resort, const double *u_free, int nmap, const int *idref2new,
                               int nd_new, double *u_new)
{
    if (nd_all < 0 || nmap < 0 || nd_new < 0 || !resort || !u_free || !u_new || (nmap > 0 && !idref2new))
        return fail(OFEA_ERR_ARG, "bad argument");
    if (nmap > nd_all) return fail(OFEA_ERR_ARG, "more map entries than nodes");
    std::fill(u_new, u_new + 3 * (size_t)nd_new, 0.0);
    for (int i = 0; i < nmap; ++i) {
        const int j = idref2new[i];
        if (j == -1) continue;
        if (j < 0 || j >= nd_new) return fail(OFEA_ERR_ARG, "idref2new entry out of range");
        const int r = resort[i];
        for (int c = 0; c < 3; ++c) u_new[3 * (size_t)j + c] = r >= 0 ? u_free[3 * (size_t)r + c] : 0.0;
    }
    return OFEA_OK;
}

// data_process::export_dis_2_vtk (src/data2vtk.cpp:47-264) and
// CA_export_dis_2_vtk (:266-427): legacy ASCII unstructured grid, default
// stream formatting (6 significant digits), tets written in node order
// (0, 2, 1, 3), hexes in VTK_VOXEL order (0, 1, 3, 2, 4, 5, 7, 6), cell types
// 10 and 11, u_x / u_y / u_z / u_magnitude and, when given, the seven nodal
// fields S_xx ... S_Mises.
int ofea_write_vtk(const char *path, int nd, const double *xyz, int ne, const int *conn, const int *etype,
                   const double *u_full, double scale, const double *node_fields)
{
    if (!path || nd < 0 || ne < 0 || (nd > 0 && (!xyz || !u_full)) || (ne > 0 && (!conn || !etype)))
        return fail(OFEA_ERR_ARG, "bad argument");
    std::ofstream f(path);
    if (!f) return fail(OFEA_ERR_IO, std::string("cannot open ") + path);
    f << "# vtk DataFile Version 3.0\n";
    f << "The density field of the optimized results\n";
    f << "ASCII\n";
    f << "DATASET UNSTRUCTURED_GRID\n\n";
    f << "POINTS\t" << nd << "\tdouble\n";
    for (int i = 0; i < nd; ++i)
        f << xyz[3 * i] + scale * u_full[3 * i] << "\t\t" << xyz[3 * i + 1] + scale * u_full[3 * i + 1] << "\t\t"
          << xyz[3 * i + 2] + scale * u_full[3 * i + 2] << "\n";
    int n4 = 0, n8 = 0;
    for (int e = 0; e < ne; ++e) {
        if (etype[e] == OFEA_C3D4) ++n4;
        else if (etype[e] == OFEA_C3D8) ++n8;
        else return fail(OFEA_ERR_ARG, "unknown element type id");
    }
    static const int tet_order[4] = {0, 2, 1, 3};
    static const int hex_order[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    f << "CELLS\t" << ne << "\t" << n8 * 9 + n4 * 5 << "\n";
    for (int e = 0; e < ne; ++e) {
        const int *cn = conn + 8 * (size_t)e;
        if (etype[e] == OFEA_C3D4) {
            f << 4;
            for (int a : tet_order) f << "\t" << cn[a] - 1;
        } else {
            f << 8;
            for (int a : hex_order) f << "\t" << cn[a] - 1;
        }
        f << "\n";
    }
    f << "CELL_TYPES\t\t" << ne << "\n";
    for (int e = 0; e < ne; ++e) f << (etype[e] == OFEA_C3D4 ? 10 : 11) << "\n";
    static const char *ucomp[3] = {"u_x", "u_y", "u_z"};
    f << "\nPOINT_DATA\t" << nd;
    for (int c = 0; c < 3; ++c) {
        f << "\nSCALARS " << ucomp[c] << " double 1\n"
          << "LOOKUP_TABLE  table" << c + 1 << "\n";
        for (int i = 0; i < nd; ++i) f << u_full[3 * i + c] << "\n";
    }
    f << "\nSCALARS u_magnitude double 1\n"
      << "LOOKUP_TABLE  table4\n";
    for (int i = 0; i < nd; ++i) {
        const double *u = u_full + 3 * (size_t)i;
        f << std::sqrt(u[2] * u[2] + u[1] * u[1] + u[0] * u[0]) << "\n";
    }
    if (node_fields) {
        static const char *names[7] = {"S_xx", "S_yy", "S_zz", "S_xy", "S_yz", "S_xz", "S_Mises"};
        for (int k = 0; k < 7; ++k) {
            f << "\nSCALARS " << names[k] << " double 1\n"
              << "LOOKUP_TABLE  table" << k + 5 << "\n";
            for (int i = 0; i < nd; ++i) f << node_fields[(size_t)k * nd + i] << "\n";
        }
    }
    f.close();
    if (!f) return fail(OFEA_ERR_IO, std::string("write failed: ") + path);
    return OFEA_OK;
}

}  // extern "C"
