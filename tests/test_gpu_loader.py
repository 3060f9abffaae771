"""The device int64 CSR loader (gpart_csr_from_int64 / _files): the same arrays and
the same error messages as the host loader, across 64 MiB chunk boundaries, from
memory and from files, in both widths."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNK = 1 << 23  # int64 entries per transfer in the loader


def path_graph(n):
    """0 - 1 - ... - (n-1): rows sorted, 2(n-1) entries."""
    deg = np.full(n, 2, np.int64)
    deg[0] = deg[-1] = 1
    indptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    indices = np.empty(indptr[-1], np.int64)
    v = np.arange(n, dtype=np.int64)
    indices[indptr[1:n]] = v[1:] - 1                      # left neighbour first
    indices[indptr[: n - 1] + (v[: n - 1] > 0)] = v[: n - 1] + 1  # then the right one
    return indptr, indices


def check_arrays(g, indptr, indices):
    off, nb, ew, vw = g.arrays()
    assert np.array_equal(off, indptr)
    assert np.array_equal(nb.astype(np.int64), indices)
    assert np.all(ew == 1) and np.all(vw == 1)
    assert len(ew) == len(indices) and len(vw) == len(indptr) - 1


@pytest.mark.parametrize("width", ["standard", "big"])
def test_memory_upload_matches_input_across_chunks(hx, width):
    n = CHUNK + 12345  # offsets span two chunks; entries span three
    indptr, indices = path_graph(n)
    g = hx.Graph(indptr, indices, width=width)
    assert g.n == n and g.m == len(indices)
    check_arrays(g, indptr, indices)


@pytest.mark.parametrize("width", ["standard", "big"])
def test_file_load_matches_input_and_host_loader(hx, width, tmp_path):
    n = CHUNK + 777
    indptr, indices = path_graph(n)
    indptr.tofile(tmp_path / "indptr.bin")
    indices.tofile(tmp_path / "indices.bin")
    g = hx.Graph.load(str(tmp_path / "indptr.bin"), str(tmp_path / "indices.bin"), width=width)
    check_arrays(g, indptr, indices)
    # the validated load takes the host path; both partition identically
    small_p, small_i = hx.generate_rmat(14, 8, 0.57, 0.19, 0.19, 3)
    np.asarray(small_p, np.int64).tofile(tmp_path / "p.bin")
    np.asarray(small_i, np.int64).tofile(tmp_path / "i.bin")
    a = hx.Graph.load(str(tmp_path / "p.bin"), str(tmp_path / "i.bin"), width=width)
    b = hx.Graph.load(str(tmp_path / "p.bin"), str(tmp_path / "i.bin"), width=width, validate=True)
    la, ra = a.partition(8, 1.03, 0)
    lb, rb = b.partition(8, 1.03, 0)
    assert np.array_equal(la, lb) and ra["final_cut"] == rb["final_cut"]


def broken_cases(n):
    indptr, indices = path_graph(n)
    cases = []
    p = indptr.copy()
    p[5] = p[4] - 1
    cases.append(("CSR offsets are not monotone", p, indices))
    if n + 1 > CHUNK:
        p = indptr.copy()
        p[CHUNK] = p[CHUNK - 1] - 1  # the first entry of the second chunk
        cases.append(("CSR offsets are not monotone", p, indices))
    p = indptr.copy()
    p[-1] += 1
    cases.append(("CSR offsets do not match indices", p, indices))
    i = indices.copy()
    i[7] = n
    cases.append(("CSR contains an invalid neighbor", indptr, i))
    i = indices.copy()
    i[-3] = -1
    cases.append(("CSR contains an invalid neighbor", indptr, i))
    return cases


@pytest.mark.parametrize("n", [1000, CHUNK + 5])
def test_errors_carry_host_loader_messages(hx, n, tmp_path):
    for message, indptr, indices in broken_cases(n):
        with pytest.raises(hx.GpartError, match=message):
            hx.Graph(indptr, indices)
        indptr.tofile(tmp_path / "p.bin")
        indices.tofile(tmp_path / "i.bin")
        with pytest.raises(hx.GpartError, match=message):
            hx.Graph.load(str(tmp_path / "p.bin"), str(tmp_path / "i.bin"))


def test_empty_and_edgeless_graphs(hx):
    g = hx.Graph(np.zeros(6, np.int64), np.zeros(0, np.int64))
    assert g.n == 5 and g.m == 0
    check_arrays(g, np.zeros(6, np.int64), np.zeros(0, np.int64))


def test_validated_upload_reports_first_failing_row(hx):
    # rows: 0:[1] 1:[0,2] 2:[1,3] 3:[2,4] 4:[3] -- a path; then break rows 2 and 4
    indptr = np.array([0, 1, 3, 5, 7, 8], np.int64)
    indices = np.array([1, 0, 2, 1, 3, 2, 4, 3], np.int64)
    w = np.ones(8, np.uint64)
    hx.Graph(indptr, indices, edge_weights=w, validate=True)
    bad_w = w.copy()
    bad_w[3] = 5  # 2 -> 1 weighs 5, 1 -> 2 weighs 1: row 1 fails first (its entry 1 -> 2)
    with pytest.raises(hx.GpartError, match="reverse edge weights differ"):
        hx.Graph(indptr, indices, edge_weights=bad_w, validate=True)
    one_way = np.array([1, 0, 2, 1, 3, 2, 4, 2], np.int64)  # 4 -> 2 without 2 -> 4, and 3 -> 4 without 4 -> 3
    with pytest.raises(hx.GpartError, match="not bidirectional"):
        hx.Graph(indptr, one_way, validate=True)
    # both faults: the lower row decides (row 1's weight mismatch before row 3's missing reverse)
    with pytest.raises(hx.GpartError, match="reverse edge weights differ"):
        hx.Graph(indptr, one_way, edge_weights=bad_w, validate=True)
