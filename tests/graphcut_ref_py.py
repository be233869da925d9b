"""An independent pure-Python restatement of the reference CPU alpha-beta swap
(src/graphcut.cc:9-279) for tiny grids, written node by node with
dictionaries: the persistent graph state (excess, height, six residues per
node), graphCutConstructGraph only for the pair's nodes, row-major
Gauss-Seidel push and relabel sweeps, ten of each per BFS global relabel,
the sink-reachable set and the keep-if-cheaper rule.  Used only to check the
oracle's C restatement (oracle/graphcut_ref.c)."""
from collections import deque

S, E, N, W_, SRC, SINK = 0, 1, 2, 3, 4, 5
STEP = {S: (0, 1), E: (1, 0), N: (0, -1), W_: (-1, 0)}
OPP = {S: N, N: S, E: W_, W_: E}


def labeling_cost(data, table, f, W, H):
    """getLabelingCost (src/grid.cc:102-125) in int."""
    total = 0
    for y in range(1, H - 1):
        for x in range(1, W - 1):
            l = f[(x, y)]
            total += int(data[l][y][x])
            total += int(table[f[(x, y - 1)]][l])   # V(l, f_N) = table[b][a] with a = l
            total += int(table[f[(x + 1, y)]][l])
    return total


def alpha_beta_ref(data, table, initial, iters, max_loops=100000):
    """data[l][y][x], table[b][a] = V(a, b), initial[y][x] -> (labels[y][x], kept moves)."""
    L, H, W = len(data), len(data[0]), len(data[0][0])
    nodes = [(x, y) for y in range(H) for x in range(W)]
    inside = lambda x, y: 0 <= x < W and 0 <= y < H
    V = lambda a, b: int(table[b][a])
    top = W * H + 2
    excess = {p: 0 for p in nodes}
    height = {p: 0 for p in nodes}
    res = {p: [0] * 6 for p in nodes}
    result = {(x, y): int(initial[y][x]) for (x, y) in nodes}
    cand = dict(result)
    best = labeling_cost(data, table, result, W, H)
    kept = 0
    success = True
    t = 0
    while t < iters and success:
        t += 1
        success = False
        for a in range(L):
            for b in range(a + 1, L):
                for (x, y) in nodes:                       # graphCutConstructGraph
                    if result[(x, y)] not in (a, b):
                        continue
                    ca = cb = 0
                    for d, (dx, dy) in STEP.items():
                        q = (x + dx, y + dy)
                        if not inside(*q):
                            res[(x, y)][d] = 0
                        elif result[q] in (a, b):
                            res[(x, y)][d] = V(a, b)
                        else:
                            res[(x, y)][d] = 0
                            ca += V(a, result[q])
                            cb += V(b, result[q])
                    excess[(x, y)] = int(data[a][y][x]) + ca
                    res[(x, y)][SINK] = int(data[b][y][x]) + cb
                    height[(x, y)] = 0
                loops = 0
                while any(excess[p] != 0 and height[p] < top for p in nodes):
                    loops += 1
                    assert loops <= max_loops
                    h = {p: 0 for p in nodes}                # graphCutGlobalRelabel
                    queue = deque()
                    for p in nodes:
                        if res[p][SINK]:
                            h[p] = 1
                            queue.append(p)
                    while queue:
                        p = queue.popleft()
                        for d, (dx, dy) in STEP.items():
                            q = (p[0] + dx, p[1] + dy)
                            if inside(*q) and res[q][OPP[d]] and h[q] == 0:
                                h[q] = h[p] + 1
                                queue.append(q)
                    for p in nodes:
                        height[p] = h[p] if h[p] else max(height[p], top)
                    for _ in range(10):
                        for p in nodes:                      # graphCutPush
                            if not (excess[p] and height[p] < top):
                                continue
                            if res[p][SINK] and height[p] == 1:
                                fl = min(res[p][SINK], excess[p])
                                excess[p] -= fl
                                res[p][SINK] -= fl
                            for d, (dx, dy) in STEP.items():
                                if res[p][d]:
                                    q = (p[0] + dx, p[1] + dy)
                                    if height[p] == height[q] + 1:
                                        fl = min(res[p][d], excess[p])
                                        excess[p] -= fl
                                        excess[q] += fl
                                        res[p][d] -= fl
                                        res[q][OPP[d]] += fl
                        for p in nodes:                      # graphCutRelabel
                            if not (excess[p] and height[p] < top):
                                continue
                            if res[p][SINK]:
                                height[p] = 1
                            else:
                                m = top
                                for d, (dx, dy) in STEP.items():
                                    if res[p][d]:
                                        m = min(m, height[(p[0] + dx, p[1] + dy)])
                                height[p] = m + 1
                reach = set()                                # graphCutFindReachable
                for p in nodes:
                    if res[p][SINK] and p not in reach:
                        reach.add(p)
                        todo = [p]
                        while todo:
                            u = todo.pop()
                            for d, (dx, dy) in STEP.items():
                                q = (u[0] + dx, u[1] + dy)
                                if inside(*q) and res[q][OPP[d]] and q not in reach:
                                    reach.add(q)
                                    todo.append(q)
                for p in nodes:                              # graphCutLabelFromFlow
                    if cand[p] in (a, b):
                        cand[p] = a if p in reach else b
                c = labeling_cost(data, table, cand, W, H)
                if c < best:
                    best = c
                    result = dict(cand)
                    success = True
                    kept += 1
                else:
                    cand = dict(result)
    return [[result[(x, y)] for x in range(W)] for y in range(H)], kept
