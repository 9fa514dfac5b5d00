"""Write ``known_answers.json``: the literal input/expected-output tables that
the reference TensorLy test-suite holds for the functions on the CP / Tucker
path.

The reference test files are parsed as text with ``ast`` (nothing is
imported or executed); only literal array values are kept.  Sources:

* ``tensorly/tests/test_core.py:31-53``  -- unfoldings of arange(24) (3,4,2)
* ``tensorly/tenalg/tests/test_khatri_rao.py:35-129`` -- three KR tables
* ``tensorly/tenalg/tests/test_n_mode_product.py:21-52,93-97`` -- mode_dot /
  multi_mode_dot tables
* ``tensorly/tests/test_cp_tensor.py:91-160`` -- cp_to_tensor tables

Run from the repository root:  python tests/golden/extract_known_answers.py
"""
import ast
import json
import os
import sys

REF = "/root/reference/tensorly"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")


def _literal_lists(path, func_name):
    """Literal arguments of ``T.tensor``/``tl.tensor`` calls, and literal
    lists assigned to ``true_res``, inside ``func_name``, in source order."""
    with open(path) as fh:
        tree = ast.parse(fh.read())
    fn = next(n for n in ast.walk(tree) if isinstance(n, ast.FunctionDef) and n.name == func_name)
    found = []
    for node in ast.walk(fn):
        if isinstance(node, ast.Call) and getattr(node.func, "attr", None) == "tensor" and node.args:
            try:
                found.append((node.lineno, node.col_offset, ast.literal_eval(node.args[0])))
            except ValueError:
                pass
        elif (isinstance(node, ast.Assign) and isinstance(node.value, ast.List)
              and any(getattr(t, "id", None) == "true_res" for t in node.targets)):
            found.append((node.lineno, node.col_offset, ast.literal_eval(node.value)))
    found.sort(key=lambda t: (t[0], t[1]))
    return [v for _, _, v in found]


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not present; known_answers.json is already committed")
    core = _literal_lists(f"{REF}/tests/test_core.py", "test_unfold")
    kr = _literal_lists(f"{REF}/tenalg/tests/test_khatri_rao.py", "test_khatri_rao")
    md = _literal_lists(f"{REF}/tenalg/tests/test_n_mode_product.py", "test_mode_dot")
    mmd = _literal_lists(f"{REF}/tenalg/tests/test_n_mode_product.py", "test_multi_mode_dot")
    cpt = _literal_lists(f"{REF}/tests/test_cp_tensor.py", "test_cp_to_tensor")
    cpw = _literal_lists(f"{REF}/tests/test_cp_tensor.py", "test_cp_to_tensor_with_weights")
    out = {
        "_source": "literal tables from the reference TensorLy 0.9.0 test-suite",
        # test_core.py: [X (as listed), unfold mode0, mode1, mode2]; X itself is arange(24).reshape(3,4,2)
        "unfold_arange24_342": {"shape": [3, 4, 2], "unfoldings": core[1:4]},
        # test_khatri_rao.py: t1, t2, true_res; A, B, true_res; (U1..U4 are aranges) true_res
        "kr_classic": {"inputs": [kr[0], kr[1]], "expected": kr[2]},
        "kr_eye": {"inputs": [kr[3], kr[4]], "expected": kr[5]},
        "kr_arange4": {"inputs_arange": [[1, 10, 3, 3], [10, 22, 4, 3], [22, 28, 2, 3], [28, 34, 2, 3]],
                        "expected": kr[6]},
        # test_n_mode_product.py::test_mode_dot
        "mode_dot_matrix": {"X": md[0], "U": md[1], "mode": 0, "expected": md[2]},
        "mode_dot_vector": {"X": md[0], "v": md[3], "mode": 1, "expected": md[4]},
        "mode_dot_vector_arange": {"X_arange": [24, [3, 4, 2]], "v_arange": 4, "mode": 1,
                                    "expected": md[5]},
        # test_multi_mode_dot: X, [u0, u1], expected scalar
        "multi_mode_dot_vectors": {"X": mmd[0], "U": [mmd[1], mmd[2]], "modes": [0, 1],
                                    "expected": mmd[3]},
        # test_cp_tensor.py
        "cp_to_tensor_arange4": {"inputs_arange": [[1, 10, 3, 3], [10, 22, 4, 3], [22, 28, 2, 3],
                                                   [28, 34, 2, 3]],
                                  "expected": cpt[0]},
        "cp_to_tensor_weights": {"A_arange": [1, 5, 2, 2], "B_arange": [5, 9, 2, 2],
                                  "weights": cpw[0], "expected": cpw[1]},
    }
    with open(OUT, "w") as fh:
        json.dump(out, fh)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
