"""Write minimal_golden.json: the option strings and process counts of
c/ch7/makefile:15-25 (runminimal_1..4) with the expected stdout of
c/ch7/output/minimal.test1..4.  Run in a tree that has the reference at
/root/reference; the JSON it writes is the committed fixture."""
import json
import os
import re

REF = "/root/reference/c/ch7"
HERE = os.path.dirname(os.path.abspath(__file__))

mk = open(os.path.join(REF, "makefile")).read()
tests = []
for line in mk.splitlines():
    m = re.search(r'testit\.sh minimal "([^"]*)" (\d+) (\d+)', line)
    if not m or line.lstrip().startswith("#"):
        continue
    opts, nprocs, num = m.group(1), int(m.group(2)), int(m.group(3))
    expected = open(os.path.join(REF, "output", "minimal.test%d" % num)).read()
    tests.append({"test": num, "options": opts, "nprocs": nprocs, "expected": expected})
tests.sort(key=lambda t: t["test"])
json.dump({"source": "c/ch7/makefile:15-25 (options, process counts) and "
                     "c/ch7/output/minimal.test1..4 (expected stdout)",
           "tests": tests},
          open(os.path.join(HERE, "minimal_golden.json"), "w"), indent=1)
print(len(tests), "tests")
