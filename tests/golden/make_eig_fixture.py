"""Extract the expected-value arrays of the reference's accessor tests
(src/test_eigValues/covarianceEigenValues_test.cpp:81-137) into a JSON fixture.

Only the numeric arrays (data) are kept, keyed by the test case that holds
them.  Run from the repository root where the reference checkout exists:
    python tests/golden/make_eig_fixture.py /root/reference
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = open(os.path.join(ref, "src", "test_eigValues", "covarianceEigenValues_test.cpp")).read()
out = {}
for m in re.finditer(r"BOOST_AUTO_TEST_CASE\((\w+)\).*?static const double res\[\] = \{([^}]*)\}",
                     src, flags=re.S):
    out[m.group(1)] = [float(x) for x in m.group(2).replace("\n", " ").split(",") if x.strip()]
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eig_accessors_daliborka.json")
with open(dst, "w") as f:
    json.dump(out, f)
print({k: len(v) for k, v in out.items()})
