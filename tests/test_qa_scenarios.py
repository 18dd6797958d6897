"""QA state machine on hand-built geometries (oracle, CPU)."""
import pytest

from oracle.oracle import Oracle
import qa_scenarios as Q


@pytest.mark.parametrize("kind,d", Q.CASES)
def test_qa_timeline_oracle(kind, d):
    ora = Oracle(Q.qa_params(), Q.qa_system(kind, d))
    got = Q.run_timeline(ora, kind, d)
    Q.check_timeline(got, Q.expected_timeline(kind, d))


@pytest.mark.parametrize("kind,d", Q.CASES)
def test_qa_contact_type_and_diffs(kind, d):
    ora = Oracle(Q.qa_params(), Q.qa_system(kind, d))
    ora.qa_update()
    cont, _ = ora.get_qa_state(0)
    ctype, d1, d2 = Q.expected_contact(kind)
    assert len(cont) == 1
    c = cont[0]
    assert (c["i1"], c["i2"]) == (1, 4)
    assert c["type"] == ctype
    assert tuple(c["diff1"]) == d1 and tuple(c["diff2"]) == d2
