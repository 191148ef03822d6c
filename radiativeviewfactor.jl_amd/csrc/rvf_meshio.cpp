// Mesh-file ingestion on the host: Gmsh .msh (format 2.2 and 4.1, ASCII and
// binary) and XML VTK unstructured grids (.vtu, XML .vtk; ascii, base64 or
// appended data, raw or zlib-compressed), producing the reference's MeshData:
// node coordinates, element families and 1-based nodes, physical-group tags
// and names.
//
// Follows load_mesh (src/MeshIO.jl:94-145) with its helpers _read_nodes,
// _read_physical_groups, _synthesize_default_group, _read_surface_elements
// and _reverse_all_normals! (src/MeshIO.jl:210-304,540-562), and the ReadVTK
// extension's _load_vtu_impl (ext/RadiativeViewFactorReadVTKExt.jl:71-157).
// The reference reads .msh through the Gmsh SDK; this reader parses the file
// formats directly (Gmsh is not needed).  Where the SDK's order is an
// implementation detail, this reader fixes one:
//   * nodes in file order (Gmsh writes node blocks in the order its
//     getNodes() returns them);
//   * groups in ascending tag order (the reference iterates a Julia Dict);
//   * within a group, entities in ascending tag order, and within an entity
//     triangle blocks before quadrangle blocks (Gmsh's storage order), each
//     in file order.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include <zlib.h>

#include "../../include/rvf.h"

int rvfi_fail(int code, const std::string& msg);

struct rvf_meshfile {
    std::vector<double> coords;          // 3 x Nn
    std::vector<int8_t> family;          // N: 0 Quad8, 1 Tri6, 2 Quad4, 3 Tri3
    std::vector<int32_t> nodes;          // 8 x N, 1-based, unused 0
    std::vector<int32_t> group;          // N
    std::map<int32_t, std::string> groups;
    int32_t flags = 0;
};

namespace {

struct MeshErr {
    std::string msg;
};
[[noreturn]] void bad(const std::string& m) { throw MeshErr{m}; }

// element families read at surface_dim (ELEM_INFO, src/MeshIO.jl:20-34):
// Gmsh type -> (family, nodes read, nodes kept)
struct Fam { int code, nread, nkeep; };
bool msh_family(int type, int dim, Fam& f) {
    if (dim == 1) {
        switch (type) {
            case 1: f = {5, 2, 2}; return true;    // Line2
            case 8: f = {4, 3, 3}; return true;    // Line3 [left, right, midpoint]
            default: return false;
        }
    }
    switch (type) {
        case 2: f = {3, 3, 3}; return true;    // Tri3
        case 3: f = {2, 4, 4}; return true;    // Quad4
        case 9: f = {1, 6, 6}; return true;    // Tri6
        case 16: f = {0, 8, 8}; return true;   // Quad8
        case 10: f = {0, 9, 8}; return true;   // Quad9: centre node dropped
        default: return false;
    }
}
// dimension of a Gmsh element type (-1: unknown)
int msh_type_dim(int type) {
    switch (type) {
        case 15: return 0;
        case 1: case 8: case 26: case 27: case 28: return 1;
        case 2: case 3: case 9: case 10: case 16: case 20: case 21: case 22: case 23: case 24: case 25: case 36:
        case 37: return 2;
        case 4: case 5: case 6: case 7: case 11: case 12: case 13: case 14: case 17: case 18: case 19: case 29:
        case 30: case 31: case 92: case 93: return 3;
        default: return -1;
    }
}
bool msh_is_tri(int type) { return type == 2 || type == 9; }

// nodes per element for every Gmsh type a binary file may carry (binary
// element data cannot be skipped without it)
int msh_nodes_of(int type) {
    static const std::map<int, int> n = {
        {1, 2},   {2, 3},   {3, 4},   {4, 4},   {5, 8},   {6, 6},   {7, 5},   {8, 3},   {9, 6},   {10, 9},
        {11, 10}, {12, 27}, {13, 18}, {14, 14}, {15, 1},  {16, 8},  {17, 20}, {18, 15}, {19, 13}, {20, 9},
        {21, 10}, {22, 12}, {23, 15}, {24, 15}, {25, 21}, {26, 4},  {27, 5},  {28, 6},  {29, 20}, {30, 35},
        {31, 56}, {36, 16}, {37, 25}, {92, 64}, {93, 125}};
    auto it = n.find(type);
    return it == n.end() ? -1 : it->second;
}

// ---------------------------------------------------------------------------
// byte cursor over the whole file
// ---------------------------------------------------------------------------
struct Cur {
    const std::string& s;
    size_t p = 0;
    explicit Cur(const std::string& str) : s(str) {}
    bool eof() const { return p >= s.size(); }
    void ws() { while (p < s.size() && std::isspace((unsigned char)s[p])) p++; }
    std::string line() {
        size_t e = s.find('\n', p);
        if (e == std::string::npos) e = s.size();
        std::string l = s.substr(p, e - p);
        p = std::min(e + 1, s.size());
        while (!l.empty() && (l.back() == '\r' || l.back() == ' ' || l.back() == '\t')) l.pop_back();
        return l;
    }
    std::string word() {
        ws();
        size_t b = p;
        while (p < s.size() && !std::isspace((unsigned char)s[p])) p++;
        if (b == p) bad("unexpected end of file");
        return s.substr(b, p - b);
    }
    // numbers are parsed in place (the buffer is NUL-terminated) and must end
    // at whitespace or the end of the file
    long long integer() {
        ws();
        const char* b = s.c_str() + p;
        char* e = nullptr;
        const long long v = std::strtoll(b, &e, 10);
        if (e == b || (*e && !std::isspace((unsigned char)*e))) bad("expected an integer, found '" + peek() + "'");
        p += (size_t)(e - b);
        return v;
    }
    double real() {
        ws();
        const char* b = s.c_str() + p;
        char* e = nullptr;
        const double v = std::strtod(b, &e);
        if (e == b || (*e && !std::isspace((unsigned char)*e))) bad("expected a number, found '" + peek() + "'");
        p += (size_t)(e - b);
        return v;
    }
    std::string peek() const { return s.substr(p, std::min<size_t>(20, s.size() - p)); }
    template <class T> T bin() {
        if (p + sizeof(T) > s.size()) bad("truncated binary data");
        T v;
        std::memcpy(&v, s.data() + p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    void expect_end(const std::string& sec) {  // "$End<sec>"
        ws();
        const std::string tag = "$End" + sec;
        if (s.compare(p, tag.size(), tag) != 0) bad("missing " + tag);
        p += tag.size();
    }
    void skip_to_end(const std::string& sec) {
        const std::string tag = "\n$End" + sec;
        size_t e = s.find(tag, p);
        if (e == std::string::npos) bad("missing $End" + sec);
        p = e + tag.size();
    }
};

struct RawElem { int type; int entity; int phys; size_t off; };  // node tags at MshModel::ntags[off]

struct MshModel {
    int dim = 2;                                                // surface_dim: 2 surfaces, 1 curves
    std::vector<long long> node_tags;
    std::vector<double> xyz;
    std::map<std::pair<int, int>, std::string> phys_names;      // (dim, tag) -> name
    std::map<int, std::set<int>> surf_phys;                     // entity of dim -> physical tags (4.1)
    std::set<int> phys2;                                        // physical tags of dim
    std::vector<RawElem> elems;                                 // elements of dim, file order
    std::vector<long long> ntags;                               // their node tags, flat
    std::set<int> surf_entities;                                // entities of dim seen
    bool v2 = false;
    // curve meshes: what _orient_line3_normals! asks Gmsh for (src/MeshIO.jl:469-599)
    std::map<int, std::set<int>> face_phys;                     // dim-2 entity -> physical tags
    std::set<int> face_entities;                                // dim-2 entities
    std::map<int, std::vector<long long>> face_nodes;           // dim-2 entity -> node tags of its elements
    std::map<int, std::vector<long long>> curve_nodes;          // dim-1 entity -> nodes classified on it
    std::map<int, std::vector<long long>> curve_elem_nodes;     // dim-1 entity -> node tags of its elements (2.2)
    std::set<long long> point_nodes;                            // nodes of point elements (2.2)
};

void read_physical_names(Cur& c, MshModel& M) {
    const long long n = c.integer();
    for (long long k = 0; k < n; k++) {
        const int dim = (int)c.integer(), tag = (int)c.integer();
        c.ws();
        std::string rest = c.line();
        if (rest.size() >= 2 && rest.front() == '"') {
            const size_t q = rest.find('"', 1);
            rest = rest.substr(1, q == std::string::npos ? std::string::npos : q - 1);
        }
        M.phys_names[{dim, tag}] = rest;
    }
    c.expect_end("PhysicalNames");
}

// ---- format 2.2 ------------------------------------------------------------
void v2_nodes(Cur& c, MshModel& M, bool binary) {
    const long long n = c.integer();
    c.line();
    M.node_tags.reserve(n);
    M.xyz.reserve(3 * n);
    for (long long k = 0; k < n; k++) {
        if (binary) {
            M.node_tags.push_back(c.bin<int32_t>());
            for (int d = 0; d < 3; d++) M.xyz.push_back(c.bin<double>());
        } else {
            M.node_tags.push_back(c.integer());
            for (int d = 0; d < 3; d++) M.xyz.push_back(c.real());
        }
    }
    c.expect_end("Nodes");
}

void v2_add(MshModel& M, int type, int phys, int ent, const long long* nt, int nn) {
    const int td = msh_type_dim(type);
    if (M.dim == 1) {
        if (td == 2) {
            M.face_entities.insert(ent);
            if (phys != 0) M.face_phys[ent].insert(phys);
            M.face_nodes[ent].insert(M.face_nodes[ent].end(), nt, nt + nn);
        } else if (td == 1) {
            M.curve_elem_nodes[ent].insert(M.curve_elem_nodes[ent].end(), nt, nt + nn);
        } else if (td == 0) {
            M.point_nodes.insert(nt, nt + nn);
        }
    }
    Fam f;
    if (!msh_family(type, M.dim, f)) return;
    if (nn < f.nread) bad("element of type " + std::to_string(type) + " with too few nodes");
    M.elems.push_back({type, ent, phys, M.ntags.size()});
    M.ntags.insert(M.ntags.end(), nt, nt + f.nread);
    if (phys != 0) M.phys2.insert(phys);  // 0: no physical group
    M.surf_entities.insert(ent);
}

void v2_elements(Cur& c, MshModel& M, bool binary) {
    const long long n = c.integer();
    c.line();
    if (!binary) {
        for (long long k = 0; k < n; k++) {
            const std::string l = c.line();
            std::vector<long long> v;
            {
                size_t q = 0;
                while (q < l.size()) {
                    while (q < l.size() && std::isspace((unsigned char)l[q])) q++;
                    if (q >= l.size()) break;
                    char* e = nullptr;
                    v.push_back(std::strtoll(l.c_str() + q, &e, 10));
                    q = (size_t)(e - l.c_str());
                    if (q < l.size() && !std::isspace((unsigned char)l[q])) bad("bad element line: " + l);
                }
            }
            if (v.size() < 3) bad("bad element line: " + l);
            const int type = (int)v[1], ntags = (int)v[2];
            if ((int)v.size() < 3 + ntags) bad("bad element line: " + l);
            const int phys = ntags > 0 ? (int)v[3] : 0, ent = ntags > 1 ? (int)v[4] : 0;
            v2_add(M, type, phys, ent, v.data() + 3 + ntags, (int)v.size() - 3 - ntags);
        }
    } else {
        long long done = 0;
        while (done < n) {
            const int type = c.bin<int32_t>(), cnt = c.bin<int32_t>(), ntags = c.bin<int32_t>();
            const int nn = msh_nodes_of(type);
            if (nn < 0) bad("unknown element type " + std::to_string(type) + " in a binary file");
            if (cnt <= 0 || ntags < 0) bad("bad binary element block");
            std::vector<long long> nt(nn);
            for (int k = 0; k < cnt; k++) {
                (void)c.bin<int32_t>();  // element number
                int phys = 0, ent = 0;
                for (int t = 0; t < ntags; t++) {
                    const int v = c.bin<int32_t>();
                    if (t == 0) phys = v;
                    if (t == 1) ent = v;
                }
                for (int a = 0; a < nn; a++) nt[a] = c.bin<int32_t>();
                v2_add(M, type, phys, ent, nt.data(), nn);
            }
            done += cnt;
        }
    }
    c.expect_end("Elements");
}

// ---- format 4.1 ------------------------------------------------------------
void v4_entities(Cur& c, MshModel& M, bool binary) {
    long long num[4];
    for (int d = 0; d < 4; d++) num[d] = binary ? (long long)c.bin<uint64_t>() : c.integer();
    for (int d = 0; d < 4; d++)
        for (long long k = 0; k < num[d]; k++) {
            const int tag = binary ? c.bin<int32_t>() : (int)c.integer();
            const int nbox = d == 0 ? 3 : 6;
            for (int b = 0; b < nbox; b++) binary ? (void)c.bin<double>() : (void)c.real();
            const long long np = binary ? (long long)c.bin<uint64_t>() : c.integer();
            for (long long q = 0; q < np; q++) {
                const int pt = binary ? c.bin<int32_t>() : (int)c.integer();
                if (d == M.dim) { M.surf_phys[tag].insert(std::abs(pt)); M.phys2.insert(std::abs(pt)); }
                if (d == 2 && M.dim == 1) M.face_phys[tag].insert(std::abs(pt));
            }
            if (d > 0) {
                const long long nb = binary ? (long long)c.bin<uint64_t>() : c.integer();
                for (long long q = 0; q < nb; q++) binary ? (void)c.bin<int32_t>() : (void)c.integer();
            }
            if (d == M.dim) M.surf_entities.insert(tag);
            if (d == 2 && M.dim == 1) M.face_entities.insert(tag);
        }
    c.expect_end("Entities");
}

void v4_nodes(Cur& c, MshModel& M, bool binary) {
    auto sz = [&]() -> long long { return binary ? (long long)c.bin<uint64_t>() : c.integer(); };
    auto i32 = [&]() -> int { return binary ? c.bin<int32_t>() : (int)c.integer(); };
    const long long nblocks = sz(), nnodes = sz();
    (void)sz(); (void)sz();
    M.node_tags.reserve(nnodes);
    M.xyz.reserve(3 * nnodes);
    for (long long b = 0; b < nblocks; b++) {
        const int edim = i32();
        const int etag = i32();
        const int param = i32();
        const long long n = sz();
        const size_t first = M.node_tags.size();
        for (long long k = 0; k < n; k++) M.node_tags.push_back(sz());
        if (M.dim == 1 && edim == 1)  // nodes classified on a curve entity (gmsh getNodes(1, ent))
            M.curve_nodes[etag].insert(M.curve_nodes[etag].end(), M.node_tags.begin() + first, M.node_tags.end());
        const int extra = param ? edim : 0;
        for (long long k = 0; k < n; k++) {
            for (int d = 0; d < 3; d++) M.xyz.push_back(binary ? c.bin<double>() : c.real());
            for (int d = 0; d < extra; d++) binary ? (void)c.bin<double>() : (void)c.real();
        }
    }
    c.expect_end("Nodes");
}

void v4_elements(Cur& c, MshModel& M, bool binary) {
    auto sz = [&]() -> long long { return binary ? (long long)c.bin<uint64_t>() : c.integer(); };
    auto i32 = [&]() -> int { return binary ? c.bin<int32_t>() : (int)c.integer(); };
    const long long nblocks = sz();
    (void)sz(); (void)sz(); (void)sz();
    for (long long b = 0; b < nblocks; b++) {
        const int edim = i32(), etag = i32(), type = i32();
        const long long n = sz();
        const int nn = msh_nodes_of(type);
        if (nn < 0) bad("unknown element type " + std::to_string(type));
        Fam f;
        const bool keep = edim == M.dim && msh_family(type, M.dim, f);
        const bool face = M.dim == 1 && edim == 2;
        std::vector<long long> nt(nn);
        for (long long k = 0; k < n; k++) {
            (void)sz();
            for (int a = 0; a < nn; a++) nt[a] = sz();
            if (keep) {
                M.elems.push_back({type, etag, 0, M.ntags.size()});
                M.ntags.insert(M.ntags.end(), nt.begin(), nt.begin() + f.nread);
            }
            if (face) M.face_nodes[etag].insert(M.face_nodes[etag].end(), nt.begin(), nt.end());
        }
        if (edim == M.dim) M.surf_entities.insert(etag);
        if (face) M.face_entities.insert(etag);
    }
    c.expect_end("Elements");
}

void read_msh(const std::string& s, MshModel& M) {
    Cur c(s);
    c.ws();
    if (c.line() != "$MeshFormat") bad("not a Gmsh mesh file ($MeshFormat missing)");
    const std::string ver = c.word();
    const int ftype = (int)c.integer(), dsize = (int)c.integer();
    c.line();
    const bool binary = ftype == 1;
    if (binary) {
        if (c.bin<int32_t>() != 1) bad("binary mesh file of the other byte order");
    }
    if (ver.rfind("2.", 0) == 0) M.v2 = true;
    else if (ver != "4.1") bad("unsupported .msh format version " + ver + " (2.2 and 4.1 are read)");
    if (binary && M.v2 && dsize != 8) bad("unsupported binary data size");
    if (binary && !M.v2 && dsize != 8) bad("unsupported binary data size (size_t must be 8 bytes)");
    c.expect_end("MeshFormat");
    bool have_nodes = false, have_elems = false;
    while (true) {
        c.ws();
        if (c.eof()) break;
        const std::string hdr = c.line();
        if (hdr.empty() || hdr[0] != '$') bad("unexpected content outside a section: " + hdr.substr(0, 40));
        const std::string sec = hdr.substr(1);
        if (sec == "PhysicalNames") read_physical_names(c, M);
        else if (sec == "Entities" && !M.v2) v4_entities(c, M, binary);
        else if (sec == "Nodes") { M.v2 ? v2_nodes(c, M, binary) : v4_nodes(c, M, binary); have_nodes = true; }
        else if (sec == "Elements") { M.v2 ? v2_elements(c, M, binary) : v4_elements(c, M, binary); have_elems = true; }
        else c.skip_to_end(sec);
    }
    if (!have_nodes || !have_elems) bad("mesh file without $Nodes or $Elements");
}

// _orient_line3_normals! (src/MeshIO.jl:469-599) with _build_curve_to_surface_centroid
// (:330-445), from the file itself instead of a Gmsh session: a curve entity's
// nodes are those classified on it (its $Nodes block in format 4.1; format
// 2.2 classifies none, so they are taken as the nodes of its line elements
// that no point element and no other curve entity uses), a surface entity's
// nodes those of its 2-D elements (entities of physical surface groups, else
// every surface entity).  Each line element whose normal at xi = 0 points away
// from the centroid of the surface next to its curve has its endpoints swapped.
template <class Lookup>
void orient_lines(const MshModel& M, const std::set<int>& curve_ents, const Lookup& lookup, rvf_meshfile& out) {
    const double* X = out.coords.data();
    auto pos_of = [&](const std::vector<long long>& tags) {
        std::set<int32_t> p;
        for (long long t : tags) { const int32_t ix = lookup(t); if (ix) p.insert(ix); }
        return p;
    };
    std::map<int, std::set<int32_t>> cnodes;
    if (!M.v2) {
        for (int ent : curve_ents) {
            auto it = M.curve_nodes.find(ent);
            cnodes[ent] = it == M.curve_nodes.end() ? std::set<int32_t>() : pos_of(it->second);
        }
    } else {
        std::map<long long, int> uses;  // node tag -> curve entities using it
        for (const auto& kv : M.curve_elem_nodes) {
            std::set<long long> u(kv.second.begin(), kv.second.end());
            for (long long t : u) uses[t]++;
        }
        for (int ent : curve_ents) {
            std::vector<long long> own;
            auto it = M.curve_elem_nodes.find(ent);
            if (it != M.curve_elem_nodes.end())
                for (long long t : it->second)
                    if (uses[t] == 1 && !M.point_nodes.count(t)) own.push_back(t);
            cnodes[ent] = pos_of(own);
        }
    }
    std::map<int32_t, int> pos_to_curve;  // first entity wins
    for (const auto& kv : cnodes)
        for (int32_t p : kv.second) pos_to_curve.emplace(p, kv.first);
    std::set<int> faces;
    for (const auto& kv : M.face_phys)
        if (!kv.second.empty()) faces.insert(kv.first);
    if (faces.empty()) faces = M.face_entities;
    std::map<int, std::set<int32_t>> fnodes;
    std::map<int, std::array<double, 3>> fcen;
    auto centroid = [&](const std::set<int32_t>& ps) {
        std::array<double, 3> c = {0.0, 0.0, 0.0};
        for (int32_t p : ps)
            for (int d = 0; d < 3; d++) c[d] += X[3 * (int64_t)(p - 1) + d];
        for (int d = 0; d < 3; d++) c[d] /= (double)ps.size();
        return c;
    };
    for (int f : faces) {
        auto it = M.face_nodes.find(f);
        if (it == M.face_nodes.end()) continue;
        std::set<int32_t> ps = pos_of(it->second);
        if (ps.empty()) continue;
        fcen[f] = centroid(ps);
        fnodes[f] = std::move(ps);
    }
    std::map<int, std::array<double, 3>> target;  // curve entity -> nearest adjacent surface centroid
    for (const auto& kv : cnodes) {
        if (kv.second.empty()) continue;
        const std::array<double, 3> cc = centroid(kv.second);
        bool found = false;
        double best = 0.0;
        for (const auto& fv : fnodes) {
            bool adj = false;
            for (int32_t p : kv.second) if (fv.second.count(p)) { adj = true; break; }
            if (!adj) continue;
            const std::array<double, 3>& sc = fcen[fv.first];
            const double d = (cc[0] - sc[0]) * (cc[0] - sc[0]) + (cc[1] - sc[1]) * (cc[1] - sc[1]) +
                             (cc[2] - sc[2]) * (cc[2] - sc[2]);
            if (!found || d < best) { best = d; target[kv.first] = sc; found = true; }
        }
    }
    for (size_t e = 0; e < out.family.size(); e++) {
        int32_t* n = &out.nodes[8 * e];
        auto it = pos_to_curve.find(n[0]);
        if (it == pos_to_curve.end()) it = pos_to_curve.find(n[1]);
        if (it == pos_to_curve.end()) { out.flags |= 8; continue; }
        auto jt = target.find(it->second);
        if (jt == target.end()) { out.flags |= 8; continue; }
        const double* x1 = X + 3 * (int64_t)(n[0] - 1);
        const double* x2 = X + 3 * (int64_t)(n[1] - 1);
        double mid[3];
        for (int d = 0; d < 3; d++) mid[d] = out.family[e] == 4 ? X[3 * (int64_t)(n[2] - 1) + d] : 0.5 * (x1[d] + x2[d]);
        const double dx = x2[0] - x1[0], dy = x2[1] - x1[1];
        const double tl = std::sqrt(dx * dx + dy * dy);
        if (tl < 2.220446049250313e-16) continue;
        const double nx = -dy / tl, ny = dx / tl;
        const double dot = (nx * (jt->second[0] - mid[0]) + ny * (jt->second[1] - mid[1])) + 0.0 * (jt->second[2] - mid[2]);
        if (dot < 0.0) { std::swap(n[0], n[1]); out.flags |= 16; }
    }
}

void build_from_msh(const MshModel& M, rvf_meshfile& out) {
    const int64_t Nn = (int64_t)M.node_tags.size();
    // node tag -> 1-based index: a dense table when tags are compact, else a map
    long long tmin = 0, tmax = -1;
    for (int64_t k = 0; k < Nn; k++) {
        tmin = k ? std::min(tmin, M.node_tags[k]) : M.node_tags[k];
        tmax = k ? std::max(tmax, M.node_tags[k]) : M.node_tags[k];
    }
    const bool dense = Nn > 0 && tmax - tmin < 4 * Nn + 1024;
    std::vector<int32_t> dense_idx;
    std::map<long long, int32_t> sparse_idx;
    if (dense) dense_idx.assign((size_t)(tmax - tmin + 1), 0);
    for (int64_t k = 0; k < Nn; k++) {
        if (dense) dense_idx[(size_t)(M.node_tags[k] - tmin)] = (int32_t)(k + 1);
        else sparse_idx[M.node_tags[k]] = (int32_t)(k + 1);
    }
    auto lookup = [&](long long tag) -> int32_t {
        if (dense) return (tag < tmin || tag > tmax) ? 0 : dense_idx[(size_t)(tag - tmin)];
        auto it = sparse_idx.find(tag);
        return it == sparse_idx.end() ? 0 : it->second;
    };
    out.coords = M.xyz;
    const std::string D = std::to_string(M.dim);
    // physical groups of dimension surface_dim (_read_physical_groups, src/MeshIO.jl:217-226)
    std::map<int, std::vector<int>> ents_of;  // physical tag -> entities of that dimension (ascending)
    std::set<int> phys = M.phys2;
    for (const auto& kv : M.phys_names)
        if (kv.first.first == M.dim) phys.insert(kv.first.second);
    if (!M.v2)
        for (const auto& kv : M.surf_phys)
            for (int p : kv.second) ents_of[p].push_back(kv.first);
    if (phys.empty()) {
        // _synthesize_default_group (src/MeshIO.jl:236-246): one group over every
        // surface entity, tagged one above the largest physical tag of any dimension
        if (M.surf_entities.empty()) bad("No entities of dimension " + D + " found in mesh; nothing to load.");
        int top = 0;
        for (const auto& kv : M.phys_names) top = std::max(top, kv.first.second);
        const int g = top + 1;
        out.groups[g] = "default";
        out.flags |= 1;
        ents_of[g] = std::vector<int>(M.surf_entities.begin(), M.surf_entities.end());
        phys.insert(g);
    } else {
        for (int g : phys) {
            auto it = M.phys_names.find({M.dim, g});
            out.groups[g] = (it == M.phys_names.end() || it->second.empty()) ? "group_" + std::to_string(g)
                                                                             : it->second;
        }
    }
    // elements per entity: triangle blocks before quadrangle blocks, file order
    std::map<int, std::vector<const RawElem*>> by_ent;
    for (const RawElem& e : M.elems) by_ent[e.entity].push_back(&e);
    for (auto& kv : by_ent)
        std::stable_sort(kv.second.begin(), kv.second.end(), [](const RawElem* a, const RawElem* b) {
            return msh_is_tri(a->type) && !msh_is_tri(b->type);
        });
    std::set<int> used_ents;  // entities of the groups read (curve orientation)
    auto push = [&](const RawElem& e, int g) {
        Fam f;
        msh_family(e.type, M.dim, f);
        used_ents.insert(e.entity);
        out.family.push_back((int8_t)f.code);
        for (int a = 0; a < 8; a++) {
            if (a < f.nkeep) {
                const long long tag = M.ntags[e.off + a];
                const int32_t ix = lookup(tag);
                if (!ix) bad("element references undefined node " + std::to_string(tag));
                out.nodes.push_back(ix);
            } else {
                out.nodes.push_back(0);
            }
        }
        out.group.push_back(g);
        if (e.type == 10) out.flags |= 2;
    };
    for (int g : phys) {
        if (M.v2 && !(out.flags & 1)) {
            // format 2.2 carries the physical tag on each element
            std::map<int, std::vector<const RawElem*>> ents;
            for (const RawElem& e : M.elems)
                if (e.phys == g) ents[e.entity].push_back(&e);
            for (auto& kv : ents) {
                std::stable_sort(kv.second.begin(), kv.second.end(), [](const RawElem* a, const RawElem* b) {
                    return msh_is_tri(a->type) && !msh_is_tri(b->type);
                });
                for (const RawElem* e : kv.second) push(*e, g);
            }
        } else if (M.v2) {
            for (auto& kv : by_ent)
                for (const RawElem* e : kv.second) push(*e, g);
        } else {
            auto it = ents_of.find(g);
            if (it == ents_of.end()) continue;
            std::vector<int> ents = it->second;
            std::sort(ents.begin(), ents.end());
            for (int ent : ents) {
                auto jt = by_ent.find(ent);
                if (jt == by_ent.end()) continue;
                for (const RawElem* e : jt->second) push(*e, g);
            }
        }
    }
    if (out.family.empty())
        bad("No supported elements found in physical groups (dimension " + D + "). Supported types (1st or 2nd "
            "order): " + (M.dim == 1 ? std::string("Line2 (1), Line3 (8)")
                                     : std::string("Tri3 (2), Quad4 (3), Tri6 (9), Quad8 (16), Quad9 (10)")));
    if (M.dim == 1) orient_lines(M, used_ents, lookup, out);
}

// ---------------------------------------------------------------------------
// XML VTK
// ---------------------------------------------------------------------------
struct XmlTag {
    std::string name;
    std::map<std::string, std::string> attr;
    size_t body = 0, end = 0;  // body start, position of the closing tag
    bool selfclose = false;
};

// next start tag at or after p named `name` (empty: any)
bool xml_find(const std::string& s, size_t p, const std::string& name, XmlTag& t, size_t limit) {
    while (true) {
        size_t lt = s.find('<', p);
        if (lt == std::string::npos || lt >= limit) return false;
        if (s.compare(lt, 2, "</") == 0 || s.compare(lt, 2, "<?") == 0 || s.compare(lt, 2, "<!") == 0) {
            p = lt + 1;
            continue;
        }
        size_t q = lt + 1, e = q;
        while (e < s.size() && !std::isspace((unsigned char)s[e]) && s[e] != '>' && s[e] != '/') e++;
        const std::string nm = s.substr(q, e - q);
        size_t gt = s.find('>', e);
        if (gt == std::string::npos) bad("unterminated XML tag");
        if (!name.empty() && nm != name) { p = gt + 1; continue; }
        t = XmlTag();
        t.name = nm;
        size_t a = e;
        while (a < gt) {
            while (a < gt && (std::isspace((unsigned char)s[a]) || s[a] == '/')) a++;
            if (a >= gt) break;
            size_t eq = s.find('=', a);
            if (eq == std::string::npos || eq > gt) break;
            std::string key = s.substr(a, eq - a);
            while (!key.empty() && std::isspace((unsigned char)key.back())) key.pop_back();
            size_t v0 = eq + 1;
            while (v0 < gt && std::isspace((unsigned char)s[v0])) v0++;
            const char qc = s[v0];
            if (qc != '"' && qc != '\'') bad("unquoted XML attribute");
            size_t v1 = s.find(qc, v0 + 1);
            if (v1 == std::string::npos || v1 > gt) bad("bad XML attribute");
            t.attr[key] = s.substr(v0 + 1, v1 - v0 - 1);
            a = v1 + 1;
        }
        t.selfclose = s[gt - 1] == '/';
        t.body = gt + 1;
        if (t.selfclose) { t.end = gt + 1; return true; }
        const std::string close = "</" + nm;
        // DataArray bodies hold no nested tags of their own name
        size_t depth = 1, scan = t.body;
        while (depth) {
            size_t nx = s.find('<', scan);
            if (nx == std::string::npos) bad("unclosed XML tag <" + nm + ">");
            if (s.compare(nx, close.size(), close) == 0 &&
                (s[nx + close.size()] == '>' || std::isspace((unsigned char)s[nx + close.size()]))) {
                if (--depth == 0) { t.end = nx; break; }
            } else if (s.compare(nx + 1, nm.size(), nm) == 0 &&
                       (std::isspace((unsigned char)s[nx + 1 + nm.size()]) || s[nx + 1 + nm.size()] == '>')) {
                size_t g2 = s.find('>', nx);
                if (g2 != std::string::npos && s[g2 - 1] != '/') depth++;
            }
            scan = nx + 1;
        }
        return true;
    }
}

// base64, possibly several padded streams back to back (VTK encodes a binary
// block's header and its data separately)
std::string b64decode(const char* p, size_t n) {
    static int8_t T[256];
    static bool init = false;
    if (!init) {
        std::memset(T, -1, sizeof T);
        const char* A = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
        for (int i = 0; i < 64; i++) T[(unsigned char)A[i]] = (int8_t)i;
        init = true;
    }
    std::string out;
    uint32_t acc = 0;
    int bits = 0;
    for (size_t k = 0; k < n; k++) {
        const unsigned char ch = (unsigned char)p[k];
        if (ch == '=') { acc = 0; bits = 0; continue; }  // end of one stream
        if (std::isspace(ch)) continue;
        if (T[ch] < 0) break;
        acc = (acc << 6) | (uint32_t)T[ch];
        bits += 6;
        if (bits >= 8) { bits -= 8; out.push_back((char)((acc >> bits) & 0xFF)); }
    }
    return out;
}

struct VtkCtx {
    const std::string& s;
    bool u64hdr = false, zlib = false;
    bool app_b64 = false;
    size_t app_start = 0;       // raw: offset of the byte after '_'
    explicit VtkCtx(const std::string& str) : s(str) {}
};

size_t type_size(const std::string& t) {
    if (t == "Int8" || t == "UInt8") return 1;
    if (t == "Int16" || t == "UInt16") return 2;
    if (t == "Int32" || t == "UInt32" || t == "Float32") return 4;
    if (t == "Int64" || t == "UInt64" || t == "Float64") return 8;
    bad("unsupported VTK data type " + t);
}

double elem_as_double(const char* b, const std::string& t) {
    if (t == "Float64") { double v; std::memcpy(&v, b, 8); return v; }
    if (t == "Float32") { float v; std::memcpy(&v, b, 4); return v; }
    if (t == "Int64") { int64_t v; std::memcpy(&v, b, 8); return (double)v; }
    if (t == "UInt64") { uint64_t v; std::memcpy(&v, b, 8); return (double)v; }
    if (t == "Int32") { int32_t v; std::memcpy(&v, b, 4); return v; }
    if (t == "UInt32") { uint32_t v; std::memcpy(&v, b, 4); return v; }
    if (t == "Int16") { int16_t v; std::memcpy(&v, b, 2); return v; }
    if (t == "UInt16") { uint16_t v; std::memcpy(&v, b, 2); return v; }
    if (t == "Int8") return (double)(int8_t)b[0];
    return (double)(uint8_t)b[0];
}

uint64_t read_hdr(const std::string& b, size_t& off, bool u64) {
    if (off + (u64 ? 8 : 4) > b.size()) bad("truncated VTK binary header");
    uint64_t v = 0;
    if (u64) std::memcpy(&v, b.data() + off, 8);
    else { uint32_t w; std::memcpy(&w, b.data() + off, 4); v = w; }
    off += u64 ? 8 : 4;
    return v;
}

// Payload of a binary block: [header][data] where the header is one byte
// count, or with zlib: block count, block size, last block size and the
// compressed size of each block.
std::string unpack_bytes(const VtkCtx& V, const std::string& b) {
    const size_t hw = V.u64hdr ? 8 : 4;
    size_t off = 0;
    if (!V.zlib) {
        const uint64_t len = read_hdr(b, off, V.u64hdr);
        if (b.size() < hw + len) bad("truncated VTK binary data");
        return b.substr(hw, len);
    }
    const uint64_t nb = read_hdr(b, off, V.u64hdr), bs = read_hdr(b, off, V.u64hdr), last = read_hdr(b, off, V.u64hdr);
    std::vector<uint64_t> cs(nb);
    for (uint64_t k = 0; k < nb; k++) cs[k] = read_hdr(b, off, V.u64hdr);
    std::string out;
    size_t pos = off;
    for (uint64_t k = 0; k < nb; k++) {
        const uint64_t raw = (k + 1 == nb && last) ? last : bs;
        if (pos + cs[k] > b.size()) bad("truncated compressed VTK block");
        std::string blk(raw, '\0');
        uLongf dl = (uLongf)raw;
        if (uncompress((Bytef*)&blk[0], &dl, (const Bytef*)b.data() + pos, (uLong)cs[k]) != Z_OK || dl != raw)
            bad("zlib: corrupt VTK block");
        out += blk;
        pos += cs[k];
    }
    return out;
}

// raw appended block at byte offset o after '_'
std::string unpack_raw(const VtkCtx& V, size_t o) {
    const size_t at = V.app_start + o;
    if (at > V.s.size()) bad("appended data offset out of range");
    return unpack_bytes(V, V.s.substr(at));
}

// values of a DataArray as doubles (integers up to 2^53 are exact)
std::vector<double> data_array(VtkCtx& V, const XmlTag& t) {
    const std::string type = t.attr.count("type") ? t.attr.at("type") : "";
    const std::string fmt = t.attr.count("format") ? t.attr.at("format") : "ascii";
    std::vector<double> v;
    if (fmt == "ascii") {
        const char* p = V.s.data() + t.body;
        const char* e = V.s.data() + t.end;
        while (p < e) {
            while (p < e && std::isspace((unsigned char)*p)) p++;
            if (p >= e) break;
            char* q = nullptr;
            v.push_back(std::strtod(p, &q));
            if (q == p) bad("bad ascii VTK data");
            p = q;
        }
        return v;
    }
    std::string raw;
    if (fmt == "binary") {
        raw = unpack_bytes(V, b64decode(V.s.data() + t.body, t.end - t.body));
    } else if (fmt == "appended") {
        if (!t.attr.count("offset")) bad("appended DataArray without an offset");
        const size_t o = (size_t)std::strtoull(t.attr.at("offset").c_str(), nullptr, 10);
        if (V.app_b64) {
            const size_t at = V.app_start + o;
            size_t lim = V.s.find('<', at);
            if (lim == std::string::npos) lim = V.s.size();
            if (at > lim) bad("appended data offset out of range");
            raw = unpack_bytes(V, b64decode(V.s.data() + at, lim - at));
        } else {
            raw = unpack_raw(V, o);
        }
    } else {
        bad("unsupported VTK DataArray format " + fmt);
    }
    const size_t w = type_size(type);
    if (raw.size() % w) bad("VTK data size not a multiple of its type");
    v.resize(raw.size() / w);
    for (size_t k = 0; k < v.size(); k++) v[k] = elem_as_double(raw.data() + k * w, type);
    return v;
}

void read_vtu(const std::string& s, int32_t surface_dim, const char* group_field, rvf_meshfile& out) {
    VtkCtx V(s);
    XmlTag root;
    if (!xml_find(s, 0, "VTKFile", root, s.size())) bad("not an XML VTK file");
    if (root.attr.count("type") && root.attr.at("type") != "UnstructuredGrid")
        bad("XML VTK file of type " + root.attr.at("type") + " (UnstructuredGrid is read)");
    if (root.attr.count("byte_order") && root.attr.at("byte_order") != "LittleEndian")
        bad("big-endian VTK data is not supported");
    V.u64hdr = root.attr.count("header_type") && root.attr.at("header_type") == "UInt64";
    if (root.attr.count("compressor")) {
        if (root.attr.at("compressor") != "vtkZLibDataCompressor")
            bad("unsupported VTK compressor " + root.attr.at("compressor"));
        V.zlib = true;
    }
    XmlTag app;
    if (xml_find(s, 0, "AppendedData", app, s.size())) {
        V.app_b64 = app.attr.count("encoding") && app.attr.at("encoding") == "base64";
        size_t u = s.find('_', app.body);
        if (u == std::string::npos) bad("AppendedData without '_'");
        V.app_start = u + 1;
    }
    XmlTag piece;
    if (!xml_find(s, root.body, "Piece", piece, root.end)) bad("VTK file without a Piece");
    const long long npts = std::atoll(piece.attr["NumberOfPoints"].c_str());
    const long long ncells = std::atoll(piece.attr["NumberOfCells"].c_str());
    XmlTag sec, da;
    if (!xml_find(s, piece.body, "Points", sec, piece.end) || !xml_find(s, sec.body, "DataArray", da, sec.end))
        bad("VTK file without Points");
    std::vector<double> P = data_array(V, da);
    if ((long long)P.size() != 3 * npts) bad("VTK Points array has the wrong size");
    out.coords = P;
    std::vector<double> conn, offs, types;
    if (!xml_find(s, piece.body, "Cells", sec, piece.end)) bad("VTK file without Cells");
    for (size_t p = sec.body; xml_find(s, p, "DataArray", da, sec.end); p = da.end) {
        const std::string nm = da.attr["Name"];
        if (nm == "connectivity") conn = data_array(V, da);
        else if (nm == "offsets") offs = data_array(V, da);
        else if (nm == "types") types = data_array(V, da);
    }
    if ((long long)types.size() != ncells || (conn.empty() && ncells)) bad("VTK Cells arrays missing or of the wrong size");
    // per-cell group array (_resolve_groups, ext/RadiativeViewFactorReadVTKExt.jl:39-61)
    std::vector<double> gv;
    bool have_g = false;
    std::vector<std::string> cands;
    if (group_field && *group_field) cands.push_back(group_field);
    else cands = {"CellEntityIds", "gmsh:physical", "RegionId", "MaterialIds", "region", "group"};
    if (xml_find(s, piece.body, "CellData", sec, piece.end)) {
        for (const std::string& want : cands) {
            for (size_t p = sec.body; xml_find(s, p, "DataArray", da, sec.end); p = da.end)
                if (da.attr["Name"] == want) {
                    std::vector<double> a = data_array(V, da);
                    if ((long long)a.size() == ncells) { gv = std::move(a); have_g = true; }
                    break;
                }
            if (have_g) break;
        }
    }
    if (!have_g && group_field && *group_field) out.flags |= 4;  // requested array absent
    // cell types -> families (VTK_FAMILY, ext/RadiativeViewFactorReadVTKExt.jl:25-32)
    for (long long k = 0; k < ncells; k++) {
        const int t = (int)types[k];
        int code = -1, nn = 0, dim = 0;
        if (t == 5) { code = 3; nn = 3; dim = 2; }
        else if (t == 22) { code = 1; nn = 6; dim = 2; }
        else if (t == 9) { code = 2; nn = 4; dim = 2; }
        else if (t == 23) { code = 0; nn = 8; dim = 2; }
        else if (t == 3) { code = 5; nn = 2; dim = 1; }
        else if (t == 21) { code = 4; nn = 3; dim = 1; }  // VTK_QUADRATIC_EDGE: ends, then midpoint
        if (code < 0 || dim != surface_dim) continue;
        long long lo, hi;
        if ((long long)offs.size() == ncells + 1) { lo = (long long)offs[k]; hi = (long long)offs[k + 1]; }
        else { lo = k == 0 ? 0 : (long long)offs[k - 1]; hi = (long long)offs[k]; }
        if (hi - lo != nn) bad("VTK cell " + std::to_string(k + 1) + " has " + std::to_string(hi - lo) +
                               " nodes, expected " + std::to_string(nn));
        if (lo < 0 || hi > (long long)conn.size()) bad("VTK connectivity out of range");
        out.family.push_back((int8_t)code);
        for (int a = 0; a < 8; a++) {
            if (a < nn) {
                const long long id = (long long)conn[lo + a];
                if (id < 0 || id >= npts) bad("VTK cell references a point out of range");
                out.nodes.push_back((int32_t)(id + 1));
            } else {
                out.nodes.push_back(0);
            }
        }
        const int g = have_g ? (int)std::llround(gv[k]) : 1;
        out.group.push_back(g);
        out.groups[g] = have_g ? "group_" + std::to_string(g) : "default";
    }
    if (out.family.empty()) bad("No VTK cells of dimension " + std::to_string(surface_dim) + " found");
}

bool is_xml_vtk(const std::string& path, const std::string& head) {
    std::string ext;
    const size_t dot = path.find_last_of('.');
    if (dot != std::string::npos) ext = path.substr(dot);
    for (char& ch : ext) ch = (char)std::tolower((unsigned char)ch);
    if (ext == ".vtu" || ext == ".pvtu" || ext == ".vts" || ext == ".vtr" || ext == ".vti" || ext == ".vtp") return true;
    if (ext != ".vtk") return false;
    return head.find("<?xml") != std::string::npos || head.find("<VTKFile") != std::string::npos;
}

// _reverse_all_normals! (src/MeshIO.jl:613-632)
void reverse_normals(rvf_meshfile& m) {
    for (size_t e = 0; e < m.family.size(); e++) {
        int32_t* n = &m.nodes[8 * e];
        if (m.family[e] >= 4) { std::swap(n[0], n[1]); continue; }  // Line2 / Line3: endpoints
        std::swap(n[0], n[2]);
        if (m.family[e] == 0) std::swap(n[4], n[6]);
        if (m.family[e] == 1) std::swap(n[3], n[5]);
    }
}

}  // namespace

extern "C" {

int rvf_meshfile_read(const char* path, int32_t surface_dim, int32_t reverse, const char* group_field,
                      rvf_meshfile** out) {
    if (!path || !out) return rvfi_fail(RVF_E_ARG, "rvf_meshfile_read: bad arguments");
    *out = nullptr;
    if (surface_dim != 1 && surface_dim != 2)
        return rvfi_fail(RVF_E_ARG, "rvf_meshfile_read: surface_dim must be 1 or 2, got " + std::to_string(surface_dim));
    FILE* f = std::fopen(path, "rb");
    if (!f) return rvfi_fail(RVF_E_ARG, std::string("Mesh file not found: ") + path);
    std::string s;
    {
        char buf[1 << 16];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
        std::fclose(f);
    }
    rvf_meshfile* m = new rvf_meshfile();
    try {
        if (is_xml_vtk(path, s.substr(0, 512))) {
            read_vtu(s, surface_dim, group_field, *m);
        } else {
            MshModel M;
            M.dim = surface_dim;
            read_msh(s, M);
            build_from_msh(M, *m);
        }
        if (reverse) reverse_normals(*m);
    } catch (const MeshErr& e) {
        delete m;
        return rvfi_fail(RVF_E_ARG, std::string(path) + ": " + e.msg);
    } catch (const std::exception& e) {
        delete m;
        return rvfi_fail(RVF_E_ARG, std::string(path) + ": " + e.what());
    }
    *out = m;
    return 0;
}

int rvf_meshfile_info(const rvf_meshfile* m, int64_t* n_nodes, int64_t* n_elems, int32_t* n_groups, int32_t* flags) {
    if (!m) return rvfi_fail(RVF_E_ARG, "rvf_meshfile_info: null mesh file");
    if (n_nodes) *n_nodes = (int64_t)(m->coords.size() / 3);
    if (n_elems) *n_elems = (int64_t)m->family.size();
    if (n_groups) *n_groups = (int32_t)m->groups.size();
    if (flags) *flags = m->flags;
    return 0;
}

int rvf_meshfile_arrays(const rvf_meshfile* m, double* coords, int8_t* family, int32_t* nodes, int32_t* group) {
    if (!m) return rvfi_fail(RVF_E_ARG, "rvf_meshfile_arrays: null mesh file");
    if (coords) std::memcpy(coords, m->coords.data(), sizeof(double) * m->coords.size());
    if (family) std::memcpy(family, m->family.data(), m->family.size());
    if (nodes) std::memcpy(nodes, m->nodes.data(), sizeof(int32_t) * m->nodes.size());
    if (group) std::memcpy(group, m->group.data(), sizeof(int32_t) * m->group.size());
    return 0;
}

int rvf_meshfile_group(const rvf_meshfile* m, int32_t k, int32_t* tag, char* name, int32_t cap) {
    if (!m || k < 0 || k >= (int32_t)m->groups.size()) return rvfi_fail(RVF_E_ARG, "rvf_meshfile_group: bad index");
    auto it = m->groups.begin();
    std::advance(it, k);
    if (tag) *tag = it->first;
    if (name && cap > 0) std::snprintf(name, (size_t)cap, "%s", it->second.c_str());
    return (int32_t)it->second.size();
}

void rvf_meshfile_free(rvf_meshfile* m) { delete m; }

}  // extern "C"
