// Host-side FV mesh service: connectivity, reverse Cuthill-McKee renumbering, FV geometry
// and boundary tags.  It stands in for the DMPlex calls the time step makes
// (VirtualSolver.hpp:386-511; TransportStep.hpp:106-120; Gradient.hpp:72-97).
//
// Geometry conventions (DMPlexGetGeometryFVM semantics, as consumed by the reference):
//   * cell volume = length (1-D) / area (2-D); centroid = segment midpoint, triangle vertex
//     mean, polygon area centroid about the vertex mean;
//   * face centroid = vertex (1-D) / edge midpoint (2-D);
//   * face normal = area-weighted (|N| = 1 in 1-D, edge length in 2-D), oriented so that
//     N . (c_R - c_L) > 0 on interior faces and N . (x_f - c_L) > 0 on boundary faces.
#include "zdm_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return -1;
}

template <typename T>
T* dup(const std::vector<T>& v) {
    T* p = static_cast<T*>(std::malloc(sizeof(T) * (v.empty() ? 1 : v.size())));
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

struct FaceRef {
    uint64_t key;
    int cell;   // input cell index
    int local;  // position in the cell's cone
};

// Reverse Cuthill-McKee over the face-adjacency graph (CSR).  Start of each component =
// unvisited vertex of minimum degree; neighbours are queued by ascending degree.
std::vector<int> rcm_order(int n, const std::vector<int>& off, const std::vector<int>& adj) {
    std::vector<int> order;
    order.reserve(n);
    std::vector<char> seen(n, 0);
    std::vector<int> deg(n);
    for (int i = 0; i < n; ++i) deg[i] = off[i + 1] - off[i];
    std::vector<int> by_deg(n);
    for (int i = 0; i < n; ++i) by_deg[i] = i;
    std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    size_t next_start = 0;
    std::vector<int> nb;
    while ((int)order.size() < n) {
        while (seen[by_deg[next_start]]) ++next_start;
        int s = by_deg[next_start];
        seen[s] = 1;
        size_t head = order.size();
        order.push_back(s);
        while (head < order.size()) {
            int v = order[head++];
            nb.clear();
            for (int k = off[v]; k < off[v + 1]; ++k)
                if (!seen[adj[k]]) nb.push_back(adj[k]);
            std::sort(nb.begin(), nb.end(), [&](int a, int b) {
                return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
            });
            for (int u : nb) {
                seen[u] = 1;
                order.push_back(u);
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// Hilbert-curve index of (x, y) on a 2^order x 2^order grid.
uint64_t hilbert_index(uint32_t order, uint32_t x, uint32_t y) {
    const uint32_t n = 1u << order;
    uint64_t d = 0;
    for (uint32_t s = n >> 1; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

// Cells ordered along a Hilbert curve through their vertex-mean centroids (2-D) or by x
// (1-D): face neighbours land in the same or adjacent 64-cell wavefronts, which keeps the
// neighbour gathers of the device kernels inside L2.
std::vector<int> hilbert_order(int dim, int ncells, int nvpc, const int* cverts,
                               const std::vector<int>& nvc, const double* xyz) {
    std::vector<double> cx(ncells), cy(ncells);
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (int c = 0; c < ncells; ++c) {
        double x = 0, y = 0;
        for (int i = 0; i < nvc[c]; ++i) {
            x += xyz[3 * (size_t)cverts[(size_t)c * nvpc + i]];
            y += xyz[3 * (size_t)cverts[(size_t)c * nvpc + i] + 1];
        }
        cx[c] = x / nvc[c];
        cy[c] = y / nvc[c];
        lo[0] = std::min(lo[0], cx[c]);
        hi[0] = std::max(hi[0], cx[c]);
        lo[1] = std::min(lo[1], cy[c]);
        hi[1] = std::max(hi[1], cy[c]);
    }
    const uint32_t order = 16;
    const double span = std::max(hi[0] - lo[0], hi[1] - lo[1]);
    const double scale = span > 0 ? ((1u << order) - 1) / span : 0.0;
    std::vector<std::pair<uint64_t, int>> key(ncells);
    for (int c = 0; c < ncells; ++c) {
        uint32_t ix = (uint32_t)((cx[c] - lo[0]) * scale);
        uint32_t iy = dim == 2 ? (uint32_t)((cy[c] - lo[1]) * scale) : 0u;
        key[c] = {dim == 2 ? hilbert_index(order, ix, iy) : (uint64_t)ix, c};
    }
    std::stable_sort(key.begin(), key.end());
    std::vector<int> perm(ncells);
    for (int c = 0; c < ncells; ++c) perm[c] = key[c].second;
    return perm;
}

uint64_t face_key(int dim, const int* fv) {
    if (dim == 1) return (uint64_t)(uint32_t)fv[0];
    uint32_t a = (uint32_t)std::min(fv[0], fv[1]), b = (uint32_t)std::max(fv[0], fv[1]);
    return ((uint64_t)a << 32) | b;
}

// Cone (faces, as vertex tuples) of a cell with k vertices.
int cell_cone(int dim, const int* cv, int k, int (*out)[2]) {
    if (dim == 1) {
        out[0][0] = cv[0];
        out[0][1] = -1;
        out[1][0] = cv[1];
        out[1][1] = -1;
        return 2;
    }
    for (int i = 0; i < k; ++i) {
        out[i][0] = cv[i];
        out[i][1] = cv[(i + 1) % k];
    }
    return k;
}

int build(int dim, int nverts, const double* xyz, int ncells, int nvpc, const int* cverts,
          int nbf, const int* bfv, const int* btag, int mark_all_tag, int renumber,
          zdm_mesh** out) {
    if (dim != 1 && dim != 2) return fail("mesh dimension must be 1 or 2");
    if (ncells <= 0 || nverts <= 0) return fail("empty mesh");
    std::vector<int> nvc(ncells);
    for (int c = 0; c < ncells; ++c) {
        int k = 0;
        while (k < nvpc && cverts[(size_t)c * nvpc + k] >= 0) ++k;
        if (dim == 1 && k != 2) return fail("1-D cells must have 2 vertices");
        if (dim == 2 && (k < 3 || k > ZDM_MESH_MAX_FACES))
            return fail("2-D cells must have 3 or 4 vertices");
        for (int i = 0; i < k; ++i) {
            int v = cverts[(size_t)c * nvpc + i];
            if (v >= nverts) return fail("cell vertex index out of range");
        }
        nvc[c] = k;
    }

    // ---- face references, grouped by key ----
    std::vector<FaceRef> refs;
    refs.reserve((size_t)ncells * (dim == 1 ? 2 : 3));
    int cone[ZDM_MESH_MAX_FACES][2];
    for (int c = 0; c < ncells; ++c) {
        int nf = cell_cone(dim, cverts + (size_t)c * nvpc, nvc[c], cone);
        for (int i = 0; i < nf; ++i) refs.push_back({face_key(dim, cone[i]), c, i});
    }
    std::sort(refs.begin(), refs.end(), [](const FaceRef& a, const FaceRef& b) {
        return a.key != b.key ? a.key < b.key : a.cell < b.cell;
    });
    struct Group {
        size_t first;
        int n;
    };
    std::vector<Group> groups;
    groups.reserve(refs.size());
    for (size_t i = 0; i < refs.size();) {
        size_t j = i + 1;
        while (j < refs.size() && refs[j].key == refs[i].key) ++j;
        if (j - i > 2) return fail("non-manifold face (more than two cells)");
        groups.push_back({i, (int)(j - i)});
        i = j;
    }

    // ---- cell adjacency (face neighbours) ----
    std::vector<int> off(ncells + 1, 0), adj;
    for (const Group& g : groups)
        if (g.n == 2) {
            off[refs[g.first].cell + 1]++;
            off[refs[g.first + 1].cell + 1]++;
        }
    for (int c = 0; c < ncells; ++c) off[c + 1] += off[c];
    adj.resize(off[ncells]);
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (const Group& g : groups)
            if (g.n == 2) {
                int a = refs[g.first].cell, b = refs[g.first + 1].cell;
                adj[fill[a]++] = b;
                adj[fill[b]++] = a;
            }
    }
    std::vector<int> perm(ncells), inv(ncells);
    if (renumber == 2)
        perm = hilbert_order(dim, ncells, nvpc, cverts, nvc, xyz);
    else if (renumber)
        perm = rcm_order(ncells, off, adj);
    else
        for (int c = 0; c < ncells; ++c) perm[c] = c;
    for (int c = 0; c < ncells; ++c) inv[perm[c]] = c;

    // ---- faces ordered by (support[0], support[1]); boundary faces after interior ones of
    //      the same support[0]; ties by the cone position in support[0] ----
    struct FaceRec {
        int s0, s1;
        int local0;
        size_t group;
    };
    std::vector<FaceRec> faces(groups.size());
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const Group& g = groups[gi];
        const FaceRef& r0 = refs[g.first];
        if (g.n == 1) {
            faces[gi] = {inv[r0.cell], -1, r0.local, gi};
        } else {
            const FaceRef& r1 = refs[g.first + 1];
            int a = inv[r0.cell], b = inv[r1.cell];
            if (a < b)
                faces[gi] = {a, b, r0.local, gi};
            else
                faces[gi] = {b, a, r1.local, gi};
        }
    }
    std::sort(faces.begin(), faces.end(), [](const FaceRec& x, const FaceRec& y) {
        if (x.s0 != y.s0) return x.s0 < y.s0;
        int xs = x.s1 < 0 ? 0x7fffffff : x.s1, ys = y.s1 < 0 ? 0x7fffffff : y.s1;
        if (xs != ys) return xs < ys;
        return x.local0 < y.local0;
    });
    const int nfaces = (int)faces.size();

    zdm_mesh* m = static_cast<zdm_mesh*>(std::calloc(1, sizeof(zdm_mesh)));
    m->dim = dim;
    m->ncells = ncells;
    m->nfaces = nfaces;
    m->nverts = nverts;
    int maxf = 0, maxv = 0;
    for (int c = 0; c < ncells; ++c) {
        maxf = std::max(maxf, dim == 1 ? 2 : nvc[c]);
        maxv = std::max(maxv, nvc[c]);
    }
    m->max_faces = maxf;
    m->max_cell_verts = maxv;

    std::vector<double> verts((size_t)nverts * 3);
    for (size_t i = 0; i < verts.size(); ++i) verts[i] = xyz[i];

    // ---- cell geometry (solver order) ----
    std::vector<double> ccen((size_t)ncells * 3, 0.0), cvol(ncells, 0.0);
    std::vector<int> cv_out((size_t)ncells * maxv, -1);
    for (int c = 0; c < ncells; ++c) {
        int o = perm[c];
        const int* cv = cverts + (size_t)o * nvpc;
        int k = nvc[o];
        for (int i = 0; i < k; ++i) cv_out[(size_t)c * maxv + i] = cv[i];
        const double* p0 = xyz + 3 * (size_t)cv[0];
        if (dim == 1) {
            const double* p1 = xyz + 3 * (size_t)cv[1];
            ccen[3 * c] = 0.5 * (p0[0] + p1[0]);
            cvol[c] = std::fabs(p1[0] - p0[0]);
        } else if (k == 3) {
            const double* p1 = xyz + 3 * (size_t)cv[1];
            const double* p2 = xyz + 3 * (size_t)cv[2];
            ccen[3 * c] = (p0[0] + p1[0] + p2[0]) / 3.0;
            ccen[3 * c + 1] = (p0[1] + p1[1] + p2[1]) / 3.0;
            double cr = (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p1[1] - p0[1]) * (p2[0] - p0[0]);
            cvol[c] = 0.5 * std::fabs(cr);
        } else {
            // polygon: shoelace about the vertex mean
            double mx = 0, my = 0;
            for (int i = 0; i < k; ++i) {
                mx += xyz[3 * (size_t)cv[i]];
                my += xyz[3 * (size_t)cv[i] + 1];
            }
            mx /= k;
            my /= k;
            double a2 = 0, cx = 0, cy = 0;
            for (int i = 0; i < k; ++i) {
                const double* a = xyz + 3 * (size_t)cv[i];
                const double* b = xyz + 3 * (size_t)cv[(i + 1) % k];
                double ax = a[0] - mx, ay = a[1] - my, bx = b[0] - mx, by = b[1] - my;
                double cr = ax * by - bx * ay;
                a2 += cr;
                cx += (ax + bx) * cr;
                cy += (ay + by) * cr;
            }
            ccen[3 * c] = mx + cx / (3.0 * a2);
            ccen[3 * c + 1] = my + cy / (3.0 * a2);
            cvol[c] = 0.5 * std::fabs(a2);
        }
    }

    // ---- tags of boundary faces ----
    std::map<uint64_t, int> tagmap;
    for (int i = 0; i < nbf; ++i) {
        int fv[2] = {bfv[(size_t)i * dim], dim == 2 ? bfv[(size_t)i * dim + 1] : -1};
        tagmap[face_key(dim, fv)] = btag[i];
    }

    // ---- face geometry ----
    std::vector<int> fcells((size_t)nfaces * 2), ftag(nfaces, -1);
    std::vector<double> fnor((size_t)nfaces * 3, 0.0), fcen((size_t)nfaces * 3, 0.0);
    for (int f = 0; f < nfaces; ++f) {
        const FaceRec& fr = faces[f];
        fcells[2 * f] = fr.s0;
        fcells[2 * f + 1] = fr.s1;
        const FaceRef& r0 = refs[groups[fr.group].first];
        int ov = r0.cell;
        const int* cv = cverts + (size_t)ov * nvpc;
        int fvv[ZDM_MESH_MAX_FACES][2];
        cell_cone(dim, cv, nvc[ov], fvv);
        const int* fv = fvv[r0.local];
        double N[3] = {0, 0, 0};
        double* X = &fcen[3 * (size_t)f];
        if (dim == 1) {
            const double* a = xyz + 3 * (size_t)fv[0];
            X[0] = a[0];
            N[0] = 1.0;
        } else {
            const double* a = xyz + 3 * (size_t)fv[0];
            const double* b = xyz + 3 * (size_t)fv[1];
            X[0] = 0.5 * (a[0] + b[0]);
            X[1] = 0.5 * (a[1] + b[1]);
            N[0] = b[1] - a[1];
            N[1] = -(b[0] - a[0]);
        }
        const double* cl = &ccen[3 * (size_t)fr.s0];
        const double* cr = fr.s1 >= 0 ? &ccen[3 * (size_t)fr.s1] : X;
        double dot = 0;
        for (int d = 0; d < dim; ++d) dot += N[d] * (cr[d] - cl[d]);
        if (dot < 0)
            for (int d = 0; d < dim; ++d) N[d] = -N[d];
        for (int d = 0; d < 3; ++d) fnor[3 * (size_t)f + d] = N[d];
        if (fr.s1 < 0) {
            auto it = tagmap.find(r0.key);
            if (it != tagmap.end())
                ftag[f] = it->second;
            else if (mark_all_tag > 0)
                ftag[f] = mark_all_tag;
        }
    }

    // ---- per-cell face lists (ascending) and neighbours ----
    std::vector<int> cnf(ncells, 0), cfaces((size_t)ncells * maxf, -1), cnbrs((size_t)ncells * maxf, -1);
    for (int f = 0; f < nfaces; ++f) {
        for (int s = 0; s < 2; ++s) {
            int c = fcells[2 * f + s];
            if (c < 0) continue;
            int k = cnf[c]++;
            if (k >= maxf) return fail("cell has more faces than expected");
            cfaces[(size_t)c * maxf + k] = f;
            cnbrs[(size_t)c * maxf + k] = fcells[2 * f + (1 - s)];
        }
    }
    // face ids were appended in ascending order because f runs ascending

    m->cell_centroid = dup(ccen);
    m->cell_volume = dup(cvol);
    m->cell_nfaces = dup(cnf);
    m->cell_faces = dup(cfaces);
    m->cell_nbrs = dup(cnbrs);
    m->cell_perm = dup(perm);
    m->cell_verts = dup(cv_out);
    m->face_cells = dup(fcells);
    m->face_normal = dup(fnor);
    m->face_centroid = dup(fcen);
    m->face_tag = dup(ftag);
    m->verts = dup(verts);
    m->nphysical = 0;
    m->physical_tag = nullptr;
    m->physical_dim = nullptr;
    m->physical_names = nullptr;
    *out = m;
    return 0;
}

}  // namespace

// errors of the other host services of this library (h5io.cpp)
void zdm_mesh_set_error(const std::string& msg) { g_err = msg; }


extern "C" {

uint64_t zdm_splitmix64_next(uint64_t* s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double zdm_splitmix64_uniform(uint64_t* s) {
    return (double)(zdm_splitmix64_next(s) >> 11) * (1.0 / 9007199254740992.0);
}

const char* zdm_mesh_last_error(void) { return g_err.c_str(); }

int zdm_mesh_from_cells(int dim, int nverts, const double* xyz, int ncells, int nvpc,
                        const int* cell_verts, int nbf, const int* bface_verts,
                        const int* bface_tags, int mark_all_tag, int renumber, zdm_mesh** out) {
    if (!out || !xyz || !cell_verts) return fail("null argument");
    try {
        return build(dim, nverts, xyz, ncells, nvpc, cell_verts, nbf, bface_verts, bface_tags,
                     mark_all_tag, renumber, out);
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

int zdm_mesh_tri_grid(int nx, int ny, double jitter, uint64_t seed, int renumber, zdm_mesh** out) {
    if (nx < 1 || ny < 1) return fail("grid needs nx, ny >= 1");
    const int nvx = nx + 1, nvy = ny + 1;
    std::vector<double> xyz((size_t)nvx * nvy * 3, 0.0);
    uint64_t st = seed;
    const double hx = 1.0 / nx, hy = 1.0 / ny;
    for (int j = 0; j < nvy; ++j)
        for (int i = 0; i < nvx; ++i) {
            size_t v = (size_t)j * nvx + i;
            double x = i * hx, y = j * hy;
            if (jitter != 0.0 && i > 0 && i < nx && j > 0 && j < ny) {
                double ux = zdm_splitmix64_uniform(&st), uy = zdm_splitmix64_uniform(&st);
                x += jitter * (2.0 * ux - 1.0) * hx;
                y += jitter * (2.0 * uy - 1.0) * hy;
            }
            xyz[3 * v] = x;
            xyz[3 * v + 1] = y;
        }
    std::vector<int> cv((size_t)nx * ny * 2 * 3);
    size_t c = 0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int sw = j * nvx + i, se = sw + 1, nw = sw + nvx, ne = nw + 1;
            cv[c++] = sw; cv[c++] = se; cv[c++] = ne;   // lower-right triangle
            cv[c++] = sw; cv[c++] = ne; cv[c++] = nw;   // upper-left triangle
        }
    return zdm_mesh_from_cells(2, nvx * nvy, xyz.data(), nx * ny * 2, 3, cv.data(), 0, nullptr,
                               nullptr, 1, renumber, out);
}

int zdm_mesh_line(int n, double x0, double x1, int renumber, zdm_mesh** out) {
    if (n < 1) return fail("line needs n >= 1");
    std::vector<double> xyz((size_t)(n + 1) * 3, 0.0);
    for (int i = 0; i <= n; ++i) xyz[3 * i] = x0 + (x1 - x0) * i / n;
    std::vector<int> cv((size_t)n * 2);
    for (int i = 0; i < n; ++i) {
        cv[2 * i] = i;
        cv[2 * i + 1] = i + 1;
    }
    int bfv[2] = {0, n}, btg[2] = {1, 2};
    return zdm_mesh_from_cells(1, n + 1, xyz.data(), n, 2, cv.data(), 2, bfv, btg, 0, renumber, out);
}

int zdm_mesh_read_msh(const char* path, int renumber, zdm_mesh** out) {
    std::ifstream in(path);
    if (!in.is_open()) return fail(std::string("cannot open mesh file: ") + path);
    std::string line;
    std::vector<int> ptag, pdim;
    std::vector<std::string> pname;
    std::map<long, int> node_index;
    std::vector<double> xyz;
    struct Elem {
        int type, phys;
        std::vector<long> nodes;
    };
    std::vector<Elem> elems;
    while (std::getline(in, line)) {
        if (line.rfind("$MeshFormat", 0) == 0) {
            std::getline(in, line);
            if (line.rfind("2.", 0) != 0) return fail("only MSH 2.x ASCII is supported");
        } else if (line.rfind("$PhysicalNames", 0) == 0) {
            int n = 0;
            std::getline(in, line);
            n = std::atoi(line.c_str());
            for (int i = 0; i < n; ++i) {
                std::getline(in, line);
                std::istringstream ss(line);
                int d, t;
                ss >> d >> t;
                std::string rest;
                std::getline(ss, rest);
                size_t a = rest.find('"'), b = rest.rfind('"');
                std::string nm = (a != std::string::npos && b > a) ? rest.substr(a + 1, b - a - 1) : rest;
                pdim.push_back(d);
                ptag.push_back(t);
                pname.push_back(nm);
            }
        } else if (line.rfind("$Nodes", 0) == 0) {
            std::getline(in, line);
            long n = std::atol(line.c_str());
            xyz.resize((size_t)n * 3);
            for (long i = 0; i < n; ++i) {
                std::getline(in, line);
                std::istringstream ss(line);
                long id;
                double x, y, z;
                ss >> id >> x >> y >> z;
                node_index[id] = (int)i;
                xyz[3 * i] = x;
                xyz[3 * i + 1] = y;
                xyz[3 * i + 2] = z;
            }
        } else if (line.rfind("$Elements", 0) == 0) {
            std::getline(in, line);
            long n = std::atol(line.c_str());
            for (long i = 0; i < n; ++i) {
                std::getline(in, line);
                std::istringstream ss(line);
                long id;
                int type, ntags;
                ss >> id >> type >> ntags;
                std::vector<int> tags(ntags);
                for (int t = 0; t < ntags; ++t) ss >> tags[t];
                int nn = type == 15 ? 1 : type == 1 ? 2 : type == 2 ? 3 : type == 3 ? 4 : -1;
                if (nn < 0) return fail("unsupported MSH element type " + std::to_string(type));
                Elem e;
                e.type = type;
                e.phys = ntags > 0 ? tags[0] : -1;
                e.nodes.resize(nn);
                for (int k = 0; k < nn; ++k) ss >> e.nodes[k];
                elems.push_back(e);
            }
        }
    }
    auto edim = [](int type) { return type == 15 ? 0 : type == 1 ? 1 : 2; };
    int dim = 0;
    for (const Elem& e : elems) dim = std::max(dim, edim(e.type));
    if (dim < 1) return fail("mesh has no cells");
    std::vector<int> cv, bfv, btag;
    int nvpc = dim == 1 ? 2 : 4;
    int ncells = 0;
    for (const Elem& e : elems) {
        int d = edim(e.type);
        if (d == dim) {
            for (int k = 0; k < nvpc; ++k)
                cv.push_back(k < (int)e.nodes.size() ? node_index.at(e.nodes[k]) : -1);
            ++ncells;
        } else if (d == dim - 1 && e.phys >= 0) {
            for (long nd : e.nodes) bfv.push_back(node_index.at(nd));
            btag.push_back(e.phys);
        }
    }
    // VirtualSolver.hpp:428-436: a mesh without physical names gets every boundary face
    // marked with value 1.
    int mark_all = pname.empty() ? 1 : 0;
    int rc = zdm_mesh_from_cells(dim, (int)(xyz.size() / 3), xyz.data(), ncells, nvpc, cv.data(),
                                 (int)btag.size(), bfv.data(), btag.data(), mark_all, renumber, out);
    if (rc) return rc;
    zdm_mesh* m = *out;
    m->nphysical = (int)pname.size();
    m->physical_tag = static_cast<int*>(std::malloc(sizeof(int) * (pname.size() + 1)));
    m->physical_dim = static_cast<int*>(std::malloc(sizeof(int) * (pname.size() + 1)));
    m->physical_names = static_cast<char*>(std::calloc(pname.size() + 1, 64));
    for (size_t i = 0; i < pname.size(); ++i) {
        m->physical_tag[i] = ptag[i];
        m->physical_dim[i] = pdim[i];
        std::strncpy(m->physical_names + 64 * i, pname[i].c_str(), 63);
    }
    return 0;
}

int zdm_mesh_concat(int n, const zdm_mesh* const* ms, int* cell_offsets, zdm_mesh** out) {
    if (n < 1 || !ms || !out) return fail("zdm_mesh_concat: bad arguments");
    const int dim = ms[0]->dim;
    size_t nc = 0, nf = 0, nv = 0;
    int maxf = 0, maxv = 0;
    for (int i = 0; i < n; ++i) {
        const zdm_mesh* a = ms[i];
        if (!a || a->dim != dim) return fail("zdm_mesh_concat: members must share the dimension");
        nc += a->ncells;
        nf += a->nfaces;
        nv += a->nverts;
        maxf = std::max(maxf, a->max_faces);
        maxv = std::max(maxv, a->max_cell_verts);
    }
    if (nc > (size_t)INT32_MAX / 4 || nf > (size_t)INT32_MAX / 2 || nv > (size_t)INT32_MAX / 3)
        return fail("zdm_mesh_concat: union too large for 32-bit indices");
    std::vector<double> ccen(nc * 3), cvol(nc), fnorm(nf * 3), fcen(nf * 3), verts(nv * 3);
    std::vector<int> cnf(nc), cfaces(nc * maxf, -1), cnbrs(nc * maxf, -1), cperm(nc),
        cverts(nc * maxv, -1), fcells(nf * 2), ftag(nf);
    size_t oc = 0, of = 0, ov = 0;
    for (int i = 0; i < n; ++i) {
        const zdm_mesh* a = ms[i];
        if (cell_offsets) cell_offsets[i] = (int)oc;
        const int MF = a->max_faces, MV = a->max_cell_verts;
        for (int c = 0; c < a->ncells; ++c) {
            const size_t g = oc + c;
            for (int d = 0; d < 3; ++d) ccen[3 * g + d] = a->cell_centroid[3 * (size_t)c + d];
            cvol[g] = a->cell_volume[c];
            cnf[g] = a->cell_nfaces[c];
            cperm[g] = (int)oc + a->cell_perm[c];
            for (int k = 0; k < MF; ++k) {
                const int f = a->cell_faces[(size_t)c * MF + k], nb = a->cell_nbrs[(size_t)c * MF + k];
                cfaces[g * maxf + k] = f < 0 ? -1 : (int)of + f;
                cnbrs[g * maxf + k] = nb < 0 ? -1 : (int)oc + nb;
            }
            for (int k = 0; k < MV; ++k) {
                const int v = a->cell_verts[(size_t)c * MV + k];
                cverts[g * maxv + k] = v < 0 ? -1 : (int)ov + v;
            }
        }
        for (int f = 0; f < a->nfaces; ++f) {
            const size_t g = of + f;
            const int L = a->face_cells[2 * f], R = a->face_cells[2 * f + 1];
            fcells[2 * g] = (int)oc + L;
            fcells[2 * g + 1] = R < 0 ? -1 : (int)oc + R;
            for (int d = 0; d < 3; ++d) {
                fnorm[3 * g + d] = a->face_normal[3 * (size_t)f + d];
                fcen[3 * g + d] = a->face_centroid[3 * (size_t)f + d];
            }
            ftag[g] = a->face_tag[f];
        }
        for (size_t v = 0; v < 3 * (size_t)a->nverts; ++v) verts[3 * ov + v] = a->verts[v];
        oc += a->ncells;
        of += a->nfaces;
        ov += a->nverts;
    }
    if (cell_offsets) cell_offsets[n] = (int)oc;
    zdm_mesh* m = static_cast<zdm_mesh*>(std::calloc(1, sizeof(zdm_mesh)));
    m->dim = dim;
    m->ncells = (int)nc;
    m->nfaces = (int)nf;
    m->max_faces = maxf;
    m->nverts = (int)nv;
    m->max_cell_verts = maxv;
    m->cell_centroid = dup(ccen);
    m->cell_volume = dup(cvol);
    m->cell_nfaces = dup(cnf);
    m->cell_faces = dup(cfaces);
    m->cell_nbrs = dup(cnbrs);
    m->cell_perm = dup(cperm);
    m->cell_verts = dup(cverts);
    m->face_cells = dup(fcells);
    m->face_normal = dup(fnorm);
    m->face_centroid = dup(fcen);
    m->face_tag = dup(ftag);
    m->verts = dup(verts);
    m->nphysical = 0;
    m->physical_tag = static_cast<int*>(std::malloc(sizeof(int)));
    m->physical_dim = static_cast<int*>(std::malloc(sizeof(int)));
    m->physical_names = static_cast<char*>(std::calloc(1, 64));
    *out = m;
    return 0;
}

void zdm_mesh_destroy(zdm_mesh* m) {
    if (!m) return;
    std::free(m->cell_centroid);
    std::free(m->cell_volume);
    std::free(m->cell_nfaces);
    std::free(m->cell_faces);
    std::free(m->cell_nbrs);
    std::free(m->cell_perm);
    std::free(m->cell_verts);
    std::free(m->face_cells);
    std::free(m->face_normal);
    std::free(m->face_centroid);
    std::free(m->face_tag);
    std::free(m->verts);
    std::free(m->physical_tag);
    std::free(m->physical_dim);
    std::free(m->physical_names);
    std::free(m);
}

}  // extern "C"
