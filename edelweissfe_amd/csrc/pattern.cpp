// Host-side construction of the system-matrix pattern and gather maps.
//
// The CSR pattern is that of edelweissfe/numerics/csrgenerator.pyx:68-77: SciPy's
// canonical csr_matrix of the VIJ pairs (dofmanager.py:543-555), i.e. for DOF row
// 3n+i every column 3m+j of every node m sharing an element with n, ascending.
// Because each element couples all DOFs of all its nodes, that pattern is the
// node adjacency expanded by 3x3, so it is stored as node-block rows (bptr/bcol)
// and expanded to CSR only on request.
//
// The element->CSR map replaces csrgenerator.pyx:81-98: instead of a COO-pair ->
// CSR-slot index (a scatter, which needs atomics or a serial loop on a GPU), every
// 3x3 block slot lists the element stiffness blocks that land on it, in element
// order (the order in which updateCSR adds them, csrgenerator.pyx:112-113), so the
// device gathers and sums each slot once, deterministically.
#include <algorithm>
#include <cstring>

#include "efe_internal.h"

bool efe_build_pattern(HostPattern& hp, int N, int E, int NN, const int32_t* conn, std::string& err) {
    hp.N = N;
    hp.E = E;
    hp.NN = NN;
    for (int64_t i = 0; i < (int64_t)E * NN; ++i) {
        if (conn[i] < 0 || conn[i] >= N) {
            err = "connectivity references a node outside [0, n_nodes)";
            return false;
        }
    }
    // node -> (element, local node) incidences, element order
    hp.nptr.assign(N + 1, 0);
    for (int64_t i = 0; i < (int64_t)E * NN; ++i) hp.nptr[conn[i] + 1]++;
    for (int n = 0; n < N; ++n) hp.nptr[n + 1] += hp.nptr[n];
    hp.nlist.resize((size_t)E * NN);
    hp.epos.resize((size_t)E * NN);
    {
        std::vector<int32_t> fill(hp.nptr.begin(), hp.nptr.end() - 1);
        for (int e = 0; e < E; ++e)
            for (int a = 0; a < NN; ++a) {
                const int32_t f = fill[conn[(int64_t)e * NN + a]]++;
                hp.nlist[f] = e * NN + a;
                hp.epos[(int64_t)a * E + e] = f;  // [a][e]: where (e, a) sits in the node-sorted list
            }
    }
    // node adjacency
    hp.bptr.assign(N + 1, 0);
    std::vector<std::vector<int32_t>> rows(N);
    for (int n = 0; n < N; ++n) {
        auto& r = rows[n];
        r.reserve(27);
        for (int32_t k = hp.nptr[n]; k < hp.nptr[n + 1]; ++k) {
            int e = hp.nlist[k] / NN;
            for (int b = 0; b < NN; ++b) r.push_back(conn[(int64_t)e * NN + b]);
        }
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        if (r.empty()) r.push_back(n);  // a node outside every element keeps its diagonal
        hp.bptr[n + 1] = hp.bptr[n] + (int32_t)r.size();
    }
    int64_t nb = hp.bptr[N];
    if (nb * 9 > INT32_MAX) {
        err = "CSR nnz exceeds the int32 index range of csrgenerator.pyx:65";
        return false;
    }
    hp.bcol.resize(nb);
    hp.diagk.resize(N);
    for (int n = 0; n < N; ++n) {
        std::copy(rows[n].begin(), rows[n].end(), hp.bcol.begin() + hp.bptr[n]);
        auto it = std::lower_bound(rows[n].begin(), rows[n].end(), n);
        hp.diagk[n] = (it != rows[n].end() && *it == n) ? (int32_t)(it - rows[n].begin()) : -1;
        std::vector<int32_t>().swap(rows[n]);
    }
    // contributions per block slot, element order
    std::vector<int32_t> slot((size_t)E * NN * NN);
    hp.coff.assign(nb + 1, 0);
    for (int e = 0; e < E; ++e) {
        const int32_t* ce = conn + (int64_t)e * NN;
        for (int a = 0; a < NN; ++a) {
            int n = ce[a];
            const int32_t* b0 = hp.bcol.data() + hp.bptr[n];
            const int32_t* b1 = hp.bcol.data() + hp.bptr[n + 1];
            for (int b = 0; b < NN; ++b) {
                const int32_t* it = std::lower_bound(b0, b1, ce[b]);
                int32_t s = (int32_t)(it - hp.bcol.data());
                slot[((size_t)e * NN + a) * NN + b] = s;
                hp.coff[s + 1]++;
            }
        }
    }
    for (int64_t s = 0; s < nb; ++s) hp.coff[s + 1] += hp.coff[s];
    // row slots: inside each window of kSortWindow nodes, rows by decreasing length (stable)
    hp.slot_node.resize(N);
    hp.node_slot.resize(N);
    for (int w0 = 0; w0 < N; w0 += kSortWindow) {
        const int w1 = std::min(N, w0 + kSortWindow);
        for (int n = w0; n < w1; ++n) hp.slot_node[n] = n;
        std::stable_sort(hp.slot_node.begin() + w0, hp.slot_node.begin() + w1, [&](int32_t a, int32_t b) {
            return hp.bptr[a + 1] - hp.bptr[a] > hp.bptr[b + 1] - hp.bptr[b];
        });
    }
    for (int p = 0; p < N; ++p) hp.node_slot[hp.slot_node[p]] = p;
    // Element scratch: contribution j of the row in slot p = 64 s + l (its contributions in
    // block order, element order inside a block: the order of updateCSR) sits at block
    // position 64 * cslice_off[s] + 64 j + l.  The element pass writes each block there; the
    // assembly, one lane per slot, reads contribution j of its 64 rows as one contiguous run
    // (no index lookups, no per-lane streams: measured on MI355X, an index per contribution
    // was latency bound and per-lane contiguous runs thrashed L2).
    {
        const int nsl = (N + kSlice - 1) / kSlice;
        hp.cslice_off.assign(nsl + 1, 0);
        for (int s = 0; s < nsl; ++s) {
            int mx = 0;
            for (int l = 0; l < kSlice; ++l) {
                const int p = s * kSlice + l;
                if (p < N) {
                    const int n = hp.slot_node[p];
                    mx = std::max(mx, hp.coff[hp.bptr[n + 1]] - hp.coff[hp.bptr[n]]);
                }
            }
            if ((int64_t)hp.cslice_off[s] + mx > INT32_MAX / kSlice) {
                err = "element scratch exceeds the int32 block index range";
                return false;
            }
            hp.cslice_off[s + 1] = hp.cslice_off[s] + mx;
        }
        hp.scratch_blocks = (int64_t)hp.cslice_off[nsl] * kSlice;
        std::vector<int32_t> fill(hp.coff.begin(), hp.coff.end() - 1);
        hp.cpos.resize((size_t)E * NN * NN);
        for (size_t k = 0; k < slot.size(); ++k) {
            const int64_t b = slot[k];
            const int32_t c = fill[b]++;
            const int64_t e = (int64_t)k / (NN * NN), a = ((int64_t)k / NN) % NN;
            const int n = conn[e * NN + a];  // row node of block (e, a, b)
            const int p = hp.node_slot[n], sl = p / kSlice, l = p % kSlice;
            hp.cpos[k] = (int32_t)(((int64_t)hp.cslice_off[sl] + (c - hp.coff[hp.bptr[n]])) * kSlice + l);
        }
    }
    // sliced-ELL layout of the block rows, in slot order
    int ns = (N + kSlice - 1) / kSlice;
    hp.slice_off.assign(ns + 1, 0);
    hp.slice_deg.assign(ns, 0);
    for (int s = 0; s < ns; ++s) {
        int d = 0;
        for (int l = 0; l < kSlice; ++l) {
            const int p = s * kSlice + l;
            if (p < N) {
                const int n = hp.slot_node[p];
                d = std::max(d, hp.bptr[n + 1] - hp.bptr[n]);
            }
        }
        hp.slice_deg[s] = d;
        hp.slice_off[s + 1] = hp.slice_off[s] + d * kSlice;
    }
    hp.padded_blocks = hp.slice_off[ns];
    hp.scol.assign(hp.padded_blocks, 0);
    for (int s = 0; s < ns; ++s) {
        for (int l = 0; l < kSlice; ++l) {
            const int p = s * kSlice + l;
            const int n = p < N ? hp.slot_node[p] : -1;
            for (int k = 0; k < hp.slice_deg[s]; ++k) {
                int32_t col = 0;
                if (n >= 0) col = (k < hp.bptr[n + 1] - hp.bptr[n]) ? hp.bcol[hp.bptr[n] + k] : n;
                hp.scol[hp.slice_off[s] + (int64_t)k * kSlice + l] = col;
            }
        }
    }
    return true;
}

// Multicolor elimination order for block ILU(0): greedy coloring of the node graph in
// natural order (no two adjacent nodes share a color), elimination order (color, node).
// The factor is stored in its own sliced-ELL layout whose rows follow that order, so
// the rows of one color are one contiguous, mutually independent range.
bool efe_build_ilu_pattern(const HostPattern& hp, IluPattern& ip, std::string& err) {
    const int N = hp.N;
    std::vector<int32_t> color(N, -1);
    std::vector<int32_t> mark;
    int ncolors = 0;
    for (int n = 0; n < N; ++n) {
        mark.assign(ncolors + 1, 0);
        for (int32_t b = hp.bptr[n]; b < hp.bptr[n + 1]; ++b) {
            int m = hp.bcol[b];
            if (m != n && color[m] >= 0) mark[color[m]] = 1;
        }
        int c = 0;
        while (mark[c]) ++c;
        color[n] = c;
        if (c + 1 > ncolors) ncolors = c + 1;
    }
    ip.ncolors = ncolors;
    ip.color_start.assign(ncolors + 1, 0);
    for (int n = 0; n < N; ++n) ip.color_start[color[n] + 1]++;
    for (int c = 0; c < ncolors; ++c) ip.color_start[c + 1] += ip.color_start[c];
    ip.perm.resize(N);   // new -> old
    ip.iperm.resize(N);  // old -> new
    {
        std::vector<int32_t> fill(ip.color_start.begin(), ip.color_start.end() - 1);
        for (int n = 0; n < N; ++n) {
            int i = fill[color[n]]++;
            ip.perm[i] = n;
            ip.iperm[n] = i;
        }
    }
    // factor rows in the new numbering, columns sorted in the new numbering
    const int ns = (N + kSlice - 1) / kSlice;
    ip.slice_off.assign(ns + 1, 0);
    ip.rowlen.assign(N, 0);
    ip.dslot.assign(N, -1);
    for (int s = 0; s < ns; ++s) {
        int d = 0;
        for (int l = 0; l < kSlice; ++l) {
            int i = s * kSlice + l;
            if (i < N) d = std::max(d, hp.bptr[ip.perm[i] + 1] - hp.bptr[ip.perm[i]]);
        }
        ip.slice_off[s + 1] = ip.slice_off[s] + d * kSlice;
    }
    const int64_t padded = ip.slice_off[ns];
    ip.scol.assign(padded, 0);
    ip.kpos.assign(padded, -1);
    std::vector<std::pair<int32_t, int32_t>> row;
    for (int i = 0; i < N; ++i) {
        const int n = ip.perm[i];
        row.clear();
        for (int32_t b = hp.bptr[n]; b < hp.bptr[n + 1]; ++b) row.emplace_back(ip.iperm[hp.bcol[b]], b - hp.bptr[n]);
        std::sort(row.begin(), row.end());
        const int s = i / kSlice, l = i % kSlice;
        const int ks = hp.node_slot[n] / kSlice, kl = hp.node_slot[n] % kSlice;
        ip.rowlen[i] = (int32_t)row.size();
        for (size_t k = 0; k < row.size(); ++k) {
            ip.scol[ip.slice_off[s] + (int64_t)k * kSlice + l] = row[k].first;
            if (row[k].first == i) ip.dslot[i] = (int32_t)k;
            // value base of the same block in the operator's sliced layout
            const int64_t kb = (int64_t)hp.slice_off[ks] * 9 + (int64_t)row[k].second * 9 * kSlice + kl;
            if (kb > INT32_MAX) {
                err = "operator too large for the 32-bit factor map";
                return false;
            }
            ip.kpos[ip.slice_off[s] + (int64_t)k * kSlice + l] = (int32_t)kb;
        }
        if (ip.dslot[i] < 0) {
            err = "a node row without its diagonal block";
            return false;
        }
    }
    return true;
}
