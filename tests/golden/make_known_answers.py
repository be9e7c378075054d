"""Writes tests/golden/known_answers.json.

Graphs come from networkx generators; the MAXCUT values for Petersen (12),
icosahedral (20) and K_{m,n} (m*n) are the known answers quoted in the
reference's scripts/maxcut_approx.py:47-58 and scripts/maxcut_exact.py:104-115.
The ER(20, 0.5, seed=42) value (the reference quotes "approximately 63",
scripts/maxcut_approx.py:60-64) is computed here by vectorised numpy brute force.
"""
import json
import os

import networkx as nx
import numpy as np


def brute_maxcut(n, edges, weights):
    masks = np.arange(1 << n, dtype=np.int64)
    best = np.zeros(masks.size)
    for (u, v), w in zip(edges, weights):
        best += w * (((masks >> u) ^ (masks >> v)) & 1)
    return float(best.max())


def entry(name, G, known=None):
    G = nx.convert_node_labels_to_integers(G)
    edges = [[int(u), int(v)] for u, v in G.edges()]
    weights = [float(G[u][v].get("weight", 1.0)) for u, v in G.edges()]
    value = known if known is not None else brute_maxcut(G.number_of_nodes(), edges, weights)
    return {"name": name, "n": G.number_of_nodes(), "edges": edges, "weights": weights, "maxcut": value}


def main():
    cases = [
        entry("petersen", nx.petersen_graph(), 12.0),
        entry("icosahedral", nx.icosahedral_graph(), 20.0),
        entry("K_8_8", nx.complete_bipartite_graph(8, 8), 64.0),
        entry("K_3_5", nx.complete_bipartite_graph(3, 5), 15.0),
        entry("er_20_0.5_seed42", nx.erdos_renyi_graph(20, 0.5, seed=42)),
    ]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(out, "w") as f:
        json.dump(cases, f, indent=0)
    for c in cases:
        print(c["name"], c["n"], len(c["edges"]), c["maxcut"])


if __name__ == "__main__":
    main()
