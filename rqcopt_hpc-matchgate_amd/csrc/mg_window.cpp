// Gate-window blocking of a gate program (see mg_window.hpp).
#include "mg_window.hpp"

#include "mg_sweep_tile.hpp"

#include <algorithm>
#include <cstdlib>
#include <set>

namespace mg {

int window_case(const WBlock& b, int W, int lo, int hi)
{
    int la = -1, lb = -1;
    for (int i = 0; i < W; i++) {
        if (b.pos[i] == lo) la = i;
        if (b.pos[i] == hi) lb = i;
    }
    if (la < 0 || lb < 0) return -1;
    if (la < lb) std::swap(la, lb);
    // (1,0)=0 (2,0)=1 (2,1)=2 (3,0)=3 (3,1)=4 (3,2)=5
    return la * (la - 1) / 2 + lb;
}

void window_blocks(const std::vector<Inst>& inst, int L, int W, int kmax, std::vector<int>* order,
                   std::vector<WBlock>* blocks)
{
    const int G = (int)inst.size();
    order->clear();
    blocks->clear();
    if (W > L || W > kWinMaxBits || W < 2) return;  // callers keep 2 <= W <= min(L, kWinMaxBits)
    // dependencies: the previous and next gate on each of a gate's two qubits
    std::vector<int> npred(G, 0);
    std::vector<int> succ(2 * G, -1);
    std::vector<int> last(L, -1);
    for (int g = 0; g < G; g++) {
        for (int q : {inst[g].lo, inst[g].hi}) {
            const int p = last[q];
            if (p >= 0) {
                // at most one successor per qubit of p
                if (inst[p].lo == q) succ[2 * p] = g;
                else succ[2 * p + 1] = g;
                npred[g]++;
            }
            last[q] = g;
        }
    }
    std::set<int> ready;
    for (int g = 0; g < G; g++)
        if (npred[g] == 0) ready.insert(g);
    auto schedule = [&](int g) {
        ready.erase(g);
        order->push_back(g);
        for (int k = 0; k < 2; k++) {
            const int s = succ[2 * g + k];
            if (s >= 0 && --npred[s] == 0) ready.insert(s);
        }
    };
    while (!ready.empty()) {
        WBlock b{};
        b.g0 = (int)order->size();
        std::vector<int> win;
        auto in_win = [&](int q) { return std::find(win.begin(), win.end(), q) != win.end(); };
        const int first = *ready.begin();
        win = {inst[first].lo, inst[first].hi};
        schedule(first);
        int n = 1;
        while (n < kmax) {
            // gates already inside the window, earliest first
            int pick = -1;
            for (int g : ready)
                if (in_win(inst[g].lo) && in_win(inst[g].hi)) {
                    pick = g;
                    break;
                }
            if (pick < 0) {
                // else widen the window by the fewest new qubits
                int best_new = 3;
                for (int g : ready) {
                    const int nw = (in_win(inst[g].lo) ? 0 : 1) + (in_win(inst[g].hi) ? 0 : 1);
                    if ((int)win.size() + nw <= W && nw < best_new) {
                        best_new = nw;
                        pick = g;
                    }
                }
                if (pick < 0) break;
                for (int q : {inst[pick].lo, inst[pick].hi})
                    if (!in_win(q)) win.push_back(q);
            }
            schedule(pick);
            n++;
        }
        // pad the window to W bits with the free positions nearest to its span
        while ((int)win.size() < W) {
            const int lo = *std::min_element(win.begin(), win.end()), hi = *std::max_element(win.begin(), win.end());
            int best = -1, bd = 1 << 30;
            for (int q = 0; q < L; q++) {
                if (in_win(q)) continue;
                const int d = q < lo ? lo - q : (q > hi ? q - hi : 0);
                if (d < bd) {
                    bd = d;
                    best = q;
                }
            }
            win.push_back(best);
        }
        std::sort(win.begin(), win.end());
        for (int i = 0; i < W; i++) b.pos[i] = win[i];
        b.g1 = (int)order->size();
        blocks->push_back(b);
    }
}

void tile_blocks(const std::vector<Inst>& inst, int L, int W, std::vector<int>* order, std::vector<TBlock>* blocks)
{
    const int G = (int)inst.size();
    order->clear();
    blocks->clear();
    if (W > L || W > kTileMaxBits || W <= kTileLowBits || L - W > 32 - kTileMaxBits) return;
    std::vector<int> npred(G, 0), succ(2 * G, -1), last(L, -1);
    for (int g = 0; g < G; g++)
        for (int q : {inst[g].lo, inst[g].hi}) {
            const int p = last[q];
            if (p >= 0) {
                if (inst[p].lo == q) succ[2 * p] = g;
                else succ[2 * p + 1] = g;
                npred[g]++;
            }
            last[q] = g;
        }
    std::set<int> ready;
    for (int g = 0; g < G; g++)
        if (npred[g] == 0) ready.insert(g);
    auto schedule = [&](int g) {
        ready.erase(g);
        order->push_back(g);
        for (int k = 0; k < 2; k++) {
            const int sc = succ[2 * g + k];
            if (sc >= 0 && --npred[sc] == 0) ready.insert(sc);
        }
    };
    while (!ready.empty()) {
        TBlock b{};
        b.g0 = (int)order->size();
        // the lowest bits always belong to the window (contiguous runs in memory)
        std::vector<int> win;
        for (int q = 0; q < kTileLowBits; q++) win.push_back(q);
        auto in_win = [&](int q) { return std::find(win.begin(), win.end(), q) != win.end(); };
        for (;;) {
            // gates inside the window first, earliest first; else widen it by the fewest bits
            int pick = -1;
            for (int g : ready)
                if (in_win(inst[g].lo) && in_win(inst[g].hi)) {
                    pick = g;
                    break;
                }
            if (pick < 0) {
                int best_new = 3;
                for (int g : ready) {
                    const int nw = (in_win(inst[g].lo) ? 0 : 1) + (in_win(inst[g].hi) ? 0 : 1);
                    if ((int)win.size() + nw <= W && nw < best_new) {
                        best_new = nw;
                        pick = g;
                    }
                }
                if (pick < 0) break;
                for (int q : {inst[pick].lo, inst[pick].hi})
                    if (!in_win(q)) win.push_back(q);
            }
            schedule(pick);
        }
        if ((int)order->size() == b.g0) break;  // no gate fits a fresh window (cannot happen for W >= 5)
        // pad to W positions with the free bits nearest to the window's top
        while ((int)win.size() < W) {
            const int hi = *std::max_element(win.begin(), win.end());
            int best = -1, bd = 1 << 30;
            for (int q = 0; q < L; q++) {
                if (in_win(q)) continue;
                const int d = q > hi ? q - hi : hi - q;
                if (d < bd) {
                    bd = d;
                    best = q;
                }
            }
            win.push_back(best);
        }
        std::sort(win.begin(), win.end());
        for (int i = 0; i < W; i++) b.pos[i] = win[i];
        int n = 0;
        for (int q = 0; q < L; q++)
            if (!in_win(q)) b.other[n++] = q;
        b.g1 = (int)order->size();
        blocks->push_back(b);
    }
}

}  // namespace mg
