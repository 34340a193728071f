"""CPU: the permuted_dense / stacked_pd oracle against the reference's known
answers (tests/golden/pd_literals.json, from tests/test_permuted_dense.h and
tests/test_stacked_pd.h): block row/col perms, block order, coalesce sources,
and values (integer-valued, so exact in any summation order)."""
import json
import os

import numpy as np
import pytest

from oracle import pd_oracle as PO

CASES = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "pd_literals.json")))["cases"]


def to_oracle(M):
    """Oracle form of a fixture operand: PD / SPD dict, or None + CSR tuple."""
    if M["kind"] == "sparse":
        return None, M
    if M["kind"] == "pd":
        return PO.pd(M["m"], M["n"], M["row_perm"], M["col_perm"], M["X"]), None
    bl = [PO.pd(b["m"], b["n"], b["row_perm"], b["col_perm"], b["X"]) for b in M["blocks"]]
    return PO.spd(M["m"], M["n"], bl), None


def dense_of(M):
    o, s = to_oracle(M)
    if o is not None:
        return PO.dense(o)
    D = np.zeros((s["m"], s["n"]))
    for r in range(s["m"]):
        D[r, s["i"][s["p"][r]:s["p"][r + 1]]] = s["x"][s["p"][r]:s["p"][r + 1]]
    return D


def oracle_result(c):
    """Oracle structure and values of one fixture case."""
    op = c["op"]
    Ao, As = to_oracle(c["A"])
    A = dense_of(c["A"])
    d = np.asarray(c["d"]) if c.get("d") is not None else None
    if op == "DA":
        S = PO.pd(Ao["m"], Ao["n"], Ao["rp"], Ao["cp"]) if "blocks" not in Ao else \
            PO.spd(Ao["m"], Ao["n"], [PO.pd(b["m"], b["n"], b["rp"], b["cp"]) for b in Ao["blocks"]])
        return PO.with_values(S, d[:, None] * A)
    if op == "ATDA":
        return PO.with_values(PO.ata(Ao), A.T @ (d[:, None] * A))
    if op == "transpose":
        return PO.with_values(PO.transpose(Ao), A.T)
    Bo, Bs = to_oracle(c["B"])
    B = dense_of(c["B"])
    a_csr = (As["p"], As["i"], As["n"]) if As else None
    if op in ("BTA", "BTDA"):
        S = PO.bta(Ao if Ao else {"blocks": []}, Bo,
                   A_csr=a_csr, B_csr=(Bs["p"], Bs["i"], Bs["n"]) if Bs else None)
        w = d if op == "BTDA" else np.ones(A.shape[0])
        return PO.with_values(S, B.T @ (w[:, None] * A))
    if op == "BA":
        S = PO.ba(Bo, Ao, B_csr=(Bs["p"], Bs["i"], Bs["m"]) if Bs else None, A_csr=a_csr)
        return PO.with_values(S, B @ A)
    raise ValueError(op)


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_pd_oracle_matches_reference_literal(c):
    got = oracle_result(c)
    exp = c["C"]
    assert (got["m"], got["n"]) == (exp["m"], exp["n"])
    assert ("blocks" in got) == (exp["kind"] == "spd")
    gb = PO.blocks(got)
    assert len(gb) == len(exp["blocks"])
    for g, e in zip(gb, exp["blocks"]):
        assert g["rp"] == e["row_perm"]
        assert g["cp"] == e["col_perm"]
        assert np.array_equal(g["X"].ravel(), np.asarray(e["X"], float))
    if "src_p" in exp:
        src = [b["src"] for b in gb]
        assert [0] + list(np.cumsum([len(s) for s in src])) == exp["src_p"]
        assert [k for s in src for k in s] == exp["src_idx"]


def test_pd_oracle_structure_covers_random_products():
    """Every nonzero of a dense product falls inside the oracle's structure
    (a dropped block or a wrong perm leaves one outside)."""
    rng = np.random.default_rng(7)
    for _ in range(40):
        m, nA, nB = 15, 9, 8

        def rand(n):
            if rng.random() < 0.5:
                rp = sorted(rng.choice(m, 6, replace=False).tolist())
                cp = sorted(rng.choice(n, 4, replace=False).tolist())
                return PO.pd(m, n, rp, cp, rng.standard_normal((6, 4)))
            rows = rng.permutation(m)[:12]
            bl = []
            for part in np.array_split(rows, 3):
                cp = sorted(rng.choice(n, 3, replace=False).tolist())
                bl.append(PO.pd(m, n, sorted(part.tolist()), cp, rng.standard_normal((len(part), 3))))
            return PO.spd(m, n, bl)

        A, B = rand(nA), rand(nB)
        S = PO.bta(A, B)
        mask = PO.dense(PO.with_values(S, np.ones((nB, nA)))) != 0
        prod = PO.dense(B).T @ PO.dense(A)
        assert not np.any(prod[~mask])
        T = PO.ata(A)
        mask = PO.dense(PO.with_values(T, np.ones((nA, nA)))) != 0
        assert not np.any((PO.dense(A).T @ PO.dense(A))[~mask])
