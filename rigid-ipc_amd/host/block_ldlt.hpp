// Sparse LDL^T of a symmetric matrix made of 6x6 blocks (one block row
// per body), the host linear solve of the Newton step (the reference uses
// Eigen::SimplicialLDLT, src/solvers/newton_solver.cpp:490-520).
//
// Elimination runs over bodies in minimum-degree order of the body graph
// (computed while eliminating, fill included); each pivot block is
// factored as L_kk D_k L_kk^T without pivoting, so an indefinite matrix
// factors as long as no pivot vanishes, as with a simplicial LDL^T.
#pragma once

#include <array>
#include <cmath>
#include <map>
#include <set>
#include <vector>

namespace rigs {

using Block = std::array<double, 36>;

class BlockLDLT {
public:
    // diag[i]: block (i,i); off: blocks (i,j) for i != j (each unordered
    // pair once; (j,i) is the transpose). Returns false on a zero or
    // non-finite pivot.
    bool factorize(int n, std::vector<Block> diag, const std::vector<std::pair<int, int>>& pairs,
                   const std::vector<Block>& off) {
        n_ = n;
        order_.clear();
        cols_.assign(n, Col{});
        std::vector<Block> D(std::move(diag));
        std::vector<std::map<int, Block>> A(n);  // A[i][j] = block (i,j), i != j
        for (size_t e = 0; e < pairs.size(); e++) {
            const int i = pairs[e].first, j = pairs[e].second;
            if (i == j) {
                add(D[i], off[e]);
                continue;
            }
            add(A[i][j], off[e]);
            add(A[j][i], transpose(off[e]));
        }
        std::vector<char> done(n, 0);
        std::set<std::pair<int, int>> queue;  // (degree, node)
        std::vector<int> deg(n);
        // Bodies without off-diagonal blocks come first in minimum-degree
        // order (degree 0, ascending) and need no elimination: factor them
        // here, and queue only the coupled ones.
        for (int i = 0; i < n; i++) {
            deg[i] = (int)A[i].size();
            if (deg[i] == 0) {
                done[i] = 1;
                order_.push_back(i);
                if (!ldlt6(D[i], cols_[i].L, cols_[i].d)) return false;
            } else {
                queue.insert({ deg[i], i });
            }
        }
        while (!queue.empty()) {
            const int k = queue.begin()->second;
            queue.erase(queue.begin());
            done[k] = 1;
            order_.push_back(k);
            Col& col = cols_[k];
            if (!ldlt6(D[k], col.L, col.d)) return false;
            // W_ik = A_ik L_kk^-T = L_ik D_k
            std::vector<std::pair<int, Block>> W;
            W.reserve(A[k].size());
            for (auto& kv : A[k]) {
                const int i = kv.first;
                Block w = A[i][k];
                solve_right_LT(col.L, w);
                W.push_back({ i, w });
            }
            for (size_t a = 0; a < W.size(); a++) {
                const int i = W[a].first;
                Block Lik;
                for (int r = 0; r < 6; r++)
                    for (int c = 0; c < 6; c++) Lik[6 * r + c] = W[a].second[6 * r + c] / col.d[c];
                col.below.push_back({ i, Lik });
                for (size_t b = 0; b < W.size(); b++) {
                    const int j = W[b].first;
                    // A_ij -= W_ik D^-1 W_jk^T
                    Block upd;
                    for (int r = 0; r < 6; r++)
                        for (int c = 0; c < 6; c++) {
                            double s = 0.0;
                            for (int p = 0; p < 6; p++) s += Lik[6 * r + p] * W[b].second[6 * c + p];
                            upd[6 * r + c] = s;
                        }
                    if (i == j) {
                        for (int e = 0; e < 36; e++) D[i][e] -= upd[e];
                    } else {
                        auto it = A[i].find(j);
                        if (it == A[i].end()) it = A[i].insert({ j, Block{} }).first;
                        for (int e = 0; e < 36; e++) it->second[e] -= upd[e];
                    }
                }
            }
            for (auto& kv : A[k]) {
                const int i = kv.first;
                A[i].erase(k);
            }
            A[k].clear();
            for (auto& w : W) {
                const int i = w.first;
                queue.erase({ deg[i], i });
                deg[i] = (int)A[i].size();
                queue.insert({ deg[i], i });
            }
        }
        return true;
    }

    void solve(const double* b, double* x) const {
        std::vector<double> y(b, b + 6 * n_);
        for (int k : order_) {
            const Col& col = cols_[k];
            double* yk = &y[6 * k];
            for (int r = 0; r < 6; r++)
                for (int p = 0; p < r; p++) yk[r] -= col.L[6 * r + p] * yk[p];
            for (const auto& e : col.below) {
                double* yi = &y[6 * e.first];
                for (int r = 0; r < 6; r++) {
                    double s = 0.0;
                    for (int p = 0; p < 6; p++) s += e.second[6 * r + p] * yk[p];
                    yi[r] -= s;
                }
            }
        }
        for (int k = 0; k < n_; k++)
            for (int r = 0; r < 6; r++) y[6 * k + r] /= cols_[k].d[r];
        for (auto it = order_.rbegin(); it != order_.rend(); ++it) {
            const int k = *it;
            const Col& col = cols_[k];
            double* yk = &y[6 * k];
            for (const auto& e : col.below) {
                const double* xi = &y[6 * e.first];
                for (int p = 0; p < 6; p++) {
                    double s = 0.0;
                    for (int r = 0; r < 6; r++) s += e.second[6 * r + p] * xi[r];
                    yk[p] -= s;
                }
            }
            for (int r = 5; r >= 0; r--)
                for (int p = r + 1; p < 6; p++) yk[r] -= col.L[6 * p + r] * yk[p];
        }
        for (int i = 0; i < 6 * n_; i++) x[i] = y[i];
    }

private:
    struct Col {
        Block L{};                 // unit lower triangle of the pivot block
        double d[6] = { 0, 0, 0, 0, 0, 0 };
        std::vector<std::pair<int, Block>> below;  // (i, L_ik)
    };
    int n_ = 0;
    std::vector<int> order_;
    std::vector<Col> cols_;

    static void add(Block& a, const Block& b) {
        for (int e = 0; e < 36; e++) a[e] += b[e];
    }
    static Block transpose(const Block& a) {
        Block t;
        for (int r = 0; r < 6; r++)
            for (int c = 0; c < 6; c++) t[6 * c + r] = a[6 * r + c];
        return t;
    }
    // M = L diag(d) L^T, L unit lower; no pivoting.
    static bool ldlt6(const Block& M, Block& L, double* d) {
        L = Block{};
        for (int j = 0; j < 6; j++) {
            double dj = M[6 * j + j];
            for (int p = 0; p < j; p++) dj -= L[6 * j + p] * L[6 * j + p] * d[p];
            if (dj == 0.0 || !std::isfinite(dj)) return false;
            d[j] = dj;
            L[6 * j + j] = 1.0;
            for (int i = j + 1; i < 6; i++) {
                double s = M[6 * i + j];
                for (int p = 0; p < j; p++) s -= L[6 * i + p] * L[6 * j + p] * d[p];
                L[6 * i + j] = s / dj;
            }
        }
        return true;
    }
    // W <- W L^-T (rows of W solve L w^T = row^T)
    static void solve_right_LT(const Block& L, Block& W) {
        for (int r = 0; r < 6; r++) {
            double* w = &W[6 * r];
            for (int c = 0; c < 6; c++)
                for (int p = 0; p < c; p++) w[c] -= L[6 * c + p] * w[p];
        }
    }
};

}  // namespace rigs
