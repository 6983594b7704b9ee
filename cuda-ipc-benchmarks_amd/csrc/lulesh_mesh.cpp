// lulesh_mesh.cpp -- reading and checking unstructured hexahedral meshes.
//
// The file format and the boundary derivation are the reference's .lmesh
// reader (LULESH/cuda/src/lulesh.cu:855-1052).  Unlike the reference, which
// reads with unchecked fscanf calls, a short or malformed file and any node
// or element index out of range are reported as errors.
#include "lulesh_mesh.h"

#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>

namespace lulesh {

namespace {

// elemBC bits, LULESH/omp_4.0/lulesh.h:60-88
constexpr int XI_M_SYMM = 0x00001, XI_P_SYMM = 0x00008, ETA_M_SYMM = 0x00040;
constexpr int ETA_P_SYMM = 0x00200, ZETA_M_SYMM = 0x01000, ZETA_P_SYMM = 0x08000;

// whitespace-separated numbers of a text file
struct Tokens {
    std::string text;
    const char *p = nullptr;

    bool next_long(long &v)
    {
        char *end = nullptr;
        errno = 0;
        v = strtol(p, &end, 10);
        if (end == p || errno) return false;
        p = end;
        return true;
    }
    bool next_int(int &v)
    {
        long l;
        if (!next_long(l) || l < INT_MIN || l > INT_MAX) return false;
        v = (int)l;
        return true;
    }
    // whether the rest of the text can still hold `n` numbers (each takes
    // at least one character and a separator): counts read from the header
    // are checked with it before anything is sized by them
    bool room_for(long long n) const
    {
        const long long left = (long long)(text.size() - (size_t)(p - text.c_str()));
        return n >= 0 && n <= (left + 1) / 2;
    }
    // fscanf's %lf conversion: strtod, correctly rounded
    bool next_double(double &v)
    {
        char *end = nullptr;
        v = strtod(p, &end);
        if (end == p) return false;
        p = end;
        return true;
    }
};

bool read_file(const char *path, std::string &out)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

bool all_in(const std::vector<int> &a, int hi)
{
    for (int v : a)
        if (v < 0 || v >= hi) return false;
    return true;
}

// a face gets the *_SYMM bit when its four nodes are all marked
// (lulesh.cu:1002-1015 and 1032-1045: corners 0-3 are the zeta- face, 4-7
// zeta+, {0,1,4,5} eta-, {2,3,6,7} eta+, {0,3,4,7} xi-, {1,2,5,6} xi+)
void face_bits(const lulesh_mesh &m, const std::vector<char> &mark, std::vector<int> &bc)
{
    for (int i = 0; i < m.numElem; ++i) {
        int mask = 0;
        for (int j = 0; j < 8; ++j) mask |= (mark[m.nodelist[8 * (size_t)i + j]] ? 1 : 0) << j;
        if ((mask & 0x0f) == 0x0f) bc[i] |= ZETA_M_SYMM;
        if ((mask & 0xf0) == 0xf0) bc[i] |= ZETA_P_SYMM;
        if ((mask & 0x33) == 0x33) bc[i] |= ETA_M_SYMM;
        if ((mask & 0xcc) == 0xcc) bc[i] |= ETA_P_SYMM;
        if ((mask & 0x99) == 0x99) bc[i] |= XI_M_SYMM;
        if ((mask & 0x66) == 0x66) bc[i] |= XI_P_SYMM;
    }
}

}  // namespace

int mesh_finish(lulesh_mesh *m)
{
    if (m->numElem < 1 || m->numNode < 1 || m->numElem >= kMaxMeshElems ||
        (long long)m->numNode * 8 > INT_MAX)
        return LULESH_EINVAL;
    if ((long long)m->x.size() != m->numNode || (long long)m->nodelist.size() != 8LL * m->numElem ||
        (long long)m->nbr.size() != 6LL * m->numElem)
        return LULESH_EINVAL;
    if (!all_in(m->nodelist, m->numNode) || !all_in(m->nbr, m->numElem) || !all_in(m->symm[0], m->numNode) ||
        !all_in(m->symm[1], m->numNode) || !all_in(m->symm[2], m->numNode) || !all_in(m->freeSurf, m->numNode))
        return LULESH_EINVAL;
    // symmetry faces: nodes on any of the three symmetry planes
    // (lulesh.cu:978-1015); the octant corner is the first element whose
    // bits are then exactly the three minus-side symmetry faces
    // (lulesh.cu:1017-1022)
    std::vector<char> mark((size_t)m->numNode, 0);
    for (int a = 0; a < 3; ++a)
        for (int n : m->symm[a]) mark[n] = 1;
    m->elemBC.assign((size_t)m->numElem, 0);
    face_bits(*m, mark, m->elemBC);
    m->octantCorner = 0;
    for (int i = 0; i < m->numElem; ++i)
        if (m->elemBC[i] == (XI_M_SYMM | ETA_M_SYMM | ZETA_M_SYMM)) {
            m->octantCorner = i;
            break;
        }
    // free-surface faces: the reference marks them with the same *_SYMM
    // bits (lulesh.cu:1024-1045), so MonoQ mirrors the element's own
    // gradient across them
    std::fill(mark.begin(), mark.end(), 0);
    for (int n : m->freeSurf) mark[n] = 1;
    face_bits(*m, mark, m->elemBC);
    return LULESH_OK;
}

}  // namespace lulesh

using namespace lulesh;

extern "C" {

static int mesh_read(const char *path, lulesh_mesh **out)
{
    Tokens t;
    if (!read_file(path, t.text)) return LULESH_EIO;
    t.p = t.text.c_str();
    lulesh_mesh *m = new lulesh_mesh();
    auto fail = [&](int rc) {
        delete m;
        return rc;
    };
    int ne, nn;
    if (!t.next_int(ne) || !t.next_int(nn)) return fail(LULESH_EIO);
    if (ne < 1 || nn < 1 || ne >= kMaxMeshElems || (long long)nn * 8 > INT_MAX)
        return fail(LULESH_EINVAL);
    // a short file fails here, before the counts size any array
    if (!t.room_for(3LL * nn + 14LL * ne + 4)) return fail(LULESH_EIO);
    m->numElem = ne;
    m->numNode = nn;
    m->x.resize(nn);
    m->y.resize(nn);
    m->z.resize(nn);
    for (int i = 0; i < nn; ++i)
        if (!t.next_double(m->x[i]) || !t.next_double(m->y[i]) || !t.next_double(m->z[i]))
            return fail(LULESH_EIO);
    m->nodelist.resize(8 * (size_t)ne);
    for (auto &v : m->nodelist)
        if (!t.next_int(v)) return fail(LULESH_EIO);
    m->nbr.resize(6 * (size_t)ne);
    for (auto &v : m->nbr)
        if (!t.next_int(v)) return fail(LULESH_EIO);
    std::vector<int> *sets[4] = {&m->symm[0], &m->symm[1], &m->symm[2], &m->freeSurf};
    for (auto *s : sets) {
        int cnt;
        if (!t.next_int(cnt)) return fail(LULESH_EIO);
        if (cnt < 0 || cnt > INT_MAX / 2) return fail(LULESH_EINVAL);
        if (!t.room_for(cnt)) return fail(LULESH_EIO);
        s->resize((size_t)cnt);
        for (auto &v : *s)
            if (!t.next_int(v)) return fail(LULESH_EIO);
    }
    const int rc = mesh_finish(m);
    if (rc) return fail(rc);
    *out = m;
    return LULESH_OK;
}

int lulesh_mesh_read(const char *path, lulesh_mesh **out)
{
    if (!path || !out) return LULESH_EINVAL;
    try {
        return mesh_read(path, out);
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

static int mesh_create(int numElem, int numNode, const double *x, const double *y, const double *z,
                       const int *nodelist, const int *nbr, const int *symmX, int nSymmX, const int *symmY,
                       int nSymmY, const int *symmZ, int nSymmZ, const int *freeSurf, int nFree,
                       lulesh_mesh **out)
{
    if (!out || numElem < 1 || numNode < 1 || numElem >= kMaxMeshElems ||
        (long long)numNode * 8 > INT_MAX || !x || !y || !z ||
        !nodelist || !nbr || nSymmX < 0 || nSymmY < 0 || nSymmZ < 0 || nFree < 0 ||
        (nSymmX && !symmX) || (nSymmY && !symmY) || (nSymmZ && !symmZ) || (nFree && !freeSurf))
        return LULESH_EINVAL;
    lulesh_mesh *m = new lulesh_mesh();
    m->numElem = numElem;
    m->numNode = numNode;
    m->x.assign(x, x + numNode);
    m->y.assign(y, y + numNode);
    m->z.assign(z, z + numNode);
    m->nodelist.assign(nodelist, nodelist + 8 * (size_t)numElem);
    m->nbr.assign(nbr, nbr + 6 * (size_t)numElem);
    if (nSymmX) m->symm[0].assign(symmX, symmX + nSymmX);
    if (nSymmY) m->symm[1].assign(symmY, symmY + nSymmY);
    if (nSymmZ) m->symm[2].assign(symmZ, symmZ + nSymmZ);
    if (nFree) m->freeSurf.assign(freeSurf, freeSurf + nFree);
    const int rc = mesh_finish(m);
    if (rc) {
        delete m;
        return rc;
    }
    *out = m;
    return LULESH_OK;
}

int lulesh_mesh_create(int numElem, int numNode, const double *x, const double *y, const double *z,
                       const int *nodelist, const int *nbr, const int *symmX, int nSymmX, const int *symmY,
                       int nSymmY, const int *symmZ, int nSymmZ, const int *freeSurf, int nFree,
                       lulesh_mesh **out)
{
    try {
        return mesh_create(numElem, numNode, x, y, z, nodelist, nbr, symmX, nSymmX, symmY, nSymmY, symmZ,
                           nSymmZ, freeSurf, nFree, out);
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

int lulesh_mesh_destroy(lulesh_mesh *m)
{
    delete m;
    return LULESH_OK;
}

int lulesh_mesh_sizes(const lulesh_mesh *m, int *numElem, int *numNode)
{
    if (!m) return LULESH_EINVAL;
    if (numElem) *numElem = m->numElem;
    if (numNode) *numNode = m->numNode;
    return LULESH_OK;
}

int lulesh_mesh_get_elem_bc(const lulesh_mesh *m, int *out, size_t n)
{
    if (!m || !out || n < (size_t)m->numElem) return LULESH_EINVAL;
    memcpy(out, m->elemBC.data(), (size_t)m->numElem * sizeof(int));
    return m->numElem;
}

int lulesh_mesh_set_elem_bc(lulesh_mesh *m, const int *bc, size_t n)
{
    if (!m || !bc || n < (size_t)m->numElem) return LULESH_EINVAL;
    m->elemBC.assign(bc, bc + m->numElem);
    return LULESH_OK;
}

int lulesh_mesh_octant_corner(const lulesh_mesh *m) { return m ? m->octantCorner : LULESH_EINVAL; }

int lulesh_mesh_set_energy(lulesh_mesh *m, double e)
{
    if (!m || !(e > 0.0)) return LULESH_EINVAL;
    m->e_deposit = e;
    return LULESH_OK;
}

}  // extern "C"
