"""Writes numc_known_answers.json: the known-answer cases numc's own matmul
tests hold (data and expected values only), restated as a fixture.

Sources (numc tree):
  tests/matmul/test_matmul_naive.c:10-160   2x3 @ 3x2 closed forms
  tests/matmul/test_matmul_naive.c:164-240  error contract
  tests/matmul/test_matmul.c:15-166         identity / ones (fills listed)
  tests/matmul/test_matmul.c:170-586        packed-vs-naive cross checks
  tests/matmul/test_matmul.c:743-801        edge-shape sweep and tolerances
  tests/reduction/test_dot_nd.c:47-72       2-D dot closed form
  tests/reduction/test_dot_nd.c:11-158      numc_dot cases (0-D, 1-D, 2-D, N-D)
  tests/reduction/test_dot.c:10-28          1-D inner product
Run: python tests/golden/make_known_answers.py
"""
import json
import os

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "numc_known_answers.json")

data = {
    # numc_dot: shapes (a, b, out), data as flat lists ("fill" v: every
    # element v), expected flat output, float32
    "dot": [
        {"case": "0-D . 1-D", "a_shape": [], "a": [2.0], "b_shape": [4], "b": [1, 2, 3, 4],
         "out_shape": [4], "expect": [2, 4, 6, 8], "source": "test_dot_nd.c:11-28"},
        {"case": "1-D . 1-D", "a_shape": [3], "a": [1, 2, 3], "b_shape": [3], "b": [4, 5, 6],
         "out_shape": [1], "expect": [32], "source": "test_dot_nd.c:30-46"},
        {"case": "1-D . 1-D", "a_shape": [4], "a": [1, 2, 3, 4], "b_shape": [4], "b": [2, 3, 4, 5],
         "out_shape": [1], "expect": [40], "source": "test_dot.c:10-28"},
        {"case": "2-D . 2-D", "a_shape": [2, 3], "a": [1, 2, 3, 4, 5, 6], "b_shape": [3, 2],
         "b": [7, 8, 9, 10, 11, 12], "out_shape": [2, 2], "expect": [58, 64, 139, 154],
         "source": "test_dot_nd.c:48-72"},
        {"case": "2-D . 1-D", "a_shape": [2, 3], "a": [1, 2, 3, 4, 5, 6], "b_shape": [3], "b": [1, 1, 1],
         "out_shape": [2], "expect": [6, 15], "source": "test_dot_nd.c:74-94"},
        {"case": "2-D . 3-D", "a_shape": [2, 3], "a": "fill 1", "b_shape": [4, 3, 5], "b": "fill 1",
         "out_shape": [2, 4, 5], "expect": "fill 3", "source": "test_dot_nd.c:96-120"},
    ],
    "closed_form": [
        # a (2x3), b (3x2), expected (2x2), dtypes
        {"a": [[1, 2, 3], [4, 5, 6]], "b": [[7, 8], [9, 10], [11, 12]],
         "expect": [[58, 64], [139, 154]],
         "dtypes": ["int16", "int32", "int64", "float32", "float64"],
         "source": "test_matmul_naive.c:37-160; test_dot_nd.c:47-72"},
        {"a": [[1, 2, 3], [4, 5, 6]], "b": [[1, 2], [3, 4], [5, 6]],
         "expect": [[22, 28], [49, 64]], "dtypes": ["int8"],
         "source": "test_matmul_naive.c:10-35"},
    ],
    "errors": [
        {"a": [2, 3], "b": [4, 2], "out": [2, 2], "dtypes": ["float32"] * 3, "code": -3,
         "case": "inner dim mismatch", "source": "test_matmul_naive.c:164-179"},
        {"a": [2, 3], "b": [3, 2], "out": [3, 3], "dtypes": ["float32"] * 3, "code": -3,
         "case": "wrong out shape", "source": "test_matmul_naive.c:181-195"},
        {"a": [6], "b": [2, 2], "out": [2, 2], "dtypes": ["float32"] * 3, "code": -3,
         "case": "1-D input", "source": "test_matmul_naive.c:197-210"},
        {"a": [2, 2], "b": [2, 2], "out": [2, 2], "dtypes": ["float32", "float64", "float32"],
         "code": -4, "case": "dtype mismatch", "source": "test_matmul_naive.c:212-224"},
        {"a": None, "b": [2, 2], "out": [2, 2], "dtypes": ["float32"] * 3, "code": -1,
         "case": "NULL a", "source": "test_matmul_naive.c:226-240"},
        {"a": [2, 2], "b": None, "out": [2, 2], "dtypes": ["float32"] * 3, "code": -1,
         "case": "NULL b", "source": "test_matmul_naive.c:226-240"},
        {"a": [2, 2], "b": [2, 2], "out": None, "dtypes": ["float32"] * 3, "code": -1,
         "case": "NULL out", "source": "test_matmul_naive.c:226-240"},
    ],
    # A @ I == A with A[i] = fill(i) (flat index), N = 64
    "identity": [
        {"dtype": "float32", "fill": "i+1", "n": 64, "source": "test_matmul.c:15-42"},
        {"dtype": "float64", "fill": "i+1", "n": 64, "source": "test_matmul.c:78-103"},
        {"dtype": "int32", "fill": "i%10+1", "n": 64, "source": "test_matmul.c:305-324"},
        {"dtype": "uint32", "fill": "i%10+1", "n": 64, "source": "test_matmul.c:354-373"},
        {"dtype": "int16", "fill": "i%10+1", "n": 64, "source": "test_matmul.c:377-396"},
        {"dtype": "uint16", "fill": "i%10+1", "n": 64, "source": "test_matmul.c:424-443"},
        {"dtype": "int64", "fill": "i%10+1", "n": 64, "source": "test_matmul.c:447-466"},
        {"dtype": "uint64", "fill": "i%10+1", "n": 64, "source": "test_matmul.c:494-513"},
        {"dtype": "int8", "fill": "i%5+1", "n": 64, "source": "test_matmul.c:517-536"},
        {"dtype": "uint8", "fill": "i%5+1", "n": 64, "source": "test_matmul.c:567-586"},
    ],
    # all-ones (M x K) @ (K x N): every entry equals K
    "ones": [
        {"dtype": "float32", "M": 2, "K": 64, "N": 2, "source": "test_matmul.c:44-74"},
        {"dtype": "float64", "M": 2, "K": 64, "N": 2, "source": "test_matmul.c:105-133"},
        {"dtype": "float32", "M": 4, "K": 128, "N": 64, "source": "test_matmul.c:137-166"},
    ],
    # A[i] = i%7+1, B[i] = i%5+1 (flat index); compared with the naive product
    "cross": [
        {"dtype": "float32", "M": 64, "K": 64, "N": 64, "tol": 1e-3, "source": "test_matmul.c:170-200"},
        {"dtype": "float64", "M": 64, "K": 64, "N": 64, "tol": 1e-10, "source": "test_matmul.c:204-236"},
        {"dtype": "float32", "M": 256, "K": 256, "N": 256, "tol": 1e-2, "source": "test_matmul.c:237-269"},
        {"dtype": "float64", "M": 256, "K": 256, "N": 256, "tol": 1e-6, "source": "test_matmul.c:270-303"},
        {"dtype": "int32", "M": 64, "K": 64, "N": 64, "tol": 0, "source": "test_matmul.c:326-352"},
        {"dtype": "int16", "M": 64, "K": 64, "N": 64, "tol": 0, "source": "test_matmul.c:398-422"},
        {"dtype": "int64", "M": 64, "K": 64, "N": 64, "tol": 0, "source": "test_matmul.c:468-492"},
        {"dtype": "int32", "M": 128, "K": 128, "N": 128, "tol": 0, "source": "test_matmul.c:654-681"},
        {"dtype": "int16", "M": 67, "K": 67, "N": 67, "tol": 0, "fill_b": "i%3+1", "fill_a": "i%5+1",
         "source": "test_matmul.c:683-708"},
    ],
    # int8 cross check with A[i] = B[i] = i%2 (sums stay below 128)
    "int8_cross": {"M": 64, "K": 64, "N": 64, "source": "test_matmul.c:538-563"},
    # edge shapes (M, K, N, dtype, tol), fills i%7+1 / i%5+1
    "edges": [
        [1, 128, 64, "float32", 1e-3], [64, 128, 1, "float32", 1e-3],
        [1, 64, 32, "float64", 1e-10], [32, 64, 1, "float64", 1e-10],
        [100, 128, 64, "float32", 1e-3], [100, 64, 32, "float64", 1e-10],
        [100, 128, 64, "int32", 0], [64, 128, 100, "float32", 1e-3],
        [32, 64, 100, "float64", 1e-10], [100, 128, 100, "float32", 1e-3],
        [50, 64, 50, "float64", 1e-10], [144, 256, 64, "float32", 1e-2],
        [148, 256, 64, "float32", 1e-2], [64, 256, 64, "float32", 1e-2],
        [64, 300, 64, "float32", 1e-2], [7, 7, 7, "float32", 1e-3],
        [7, 7, 7, "int32", 0], [97, 101, 103, "float32", 1e-2],
        [97, 101, 103, "float64", 1e-8], [97, 101, 103, "int32", 0],
        [37, 41, 43, "int16", 0],
    ],
    "edges_source": "test_matmul.c:743-801",
}

if __name__ == "__main__":
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(OUT)
