"""Host side of the Python mirror (cbm/api.py): ABM options map to the model
descriptor, user agent parameters to their fields and to the names a rule
body sees.  CPU only: nothing here touches the device."""
import pytest

import cbm._abi as A
from cbm.api import ABM, CBMIntegrators, CBMModels, CBMNeighbors, CBMRules, Community

BODY = "struct UserRule { __device__ UserRule(const double*, const double*) {} " \
       "__device__ void operator()(RuleAgent& a) const {} };"


def test_rule_body_and_user_parameters():
    abm = ABM(2, agentRule=(BODY, [0.5, 0, 0, 0]), agent={"age": float, "gen": float})
    m = abm.model_desc(0.1)
    assert m.rule == A.RULE_CUSTOM and m.rule_p[0] == 0.5
    assert abm.user_fields == {"age": A.F_U0, "gen": A.F_U1}
    com = Community(abm, N=3, dt=0.1, simBox=[[0, 1], [0, 1]])
    assert set(com.params) >= {"x", "y", "id", "vol", "growth", "vbirth", "lf", "age", "gen"}
    com.age = 2.0
    assert com.age.tolist() == [2.0, 2.0, 2.0]
    with pytest.raises(KeyError):
        com["speed"] = 1.0


@pytest.mark.parametrize("agent", [{"x": float}, {"a b": float}, {"n": int},
                                   {"a": float, "b": float, "c": float, "d": float, "e": float}])
def test_bad_user_parameters_are_refused(agent):
    with pytest.raises(ValueError):
        ABM(2, agentRule=BODY, agent=agent)


def test_user_parameters_need_a_rule_body():
    with pytest.raises(ValueError):
        ABM(2, agentRule=CBMRules.death(0.1), agent={"age": float})


def test_medium_and_noise_options():
    abm = ABM(2, agentODE=CBMModels.linearRepulsion(1.0, 2.0), neighborsAlg=CBMNeighbors.CellLinked(cellEdge=2.0),
              agentAlg=CBMIntegrators.EM(), agentSDE="struct UserNoise {};",
              medium={"NMedium": [16, 8], "D": 0.5, "consume": 0.1, "c0": 1.0, "mediumAlg": "DGADI"})
    m = abm.model_desc(0.01)
    assert m.medium_alg == A.MEDIUM_DGADI and list(m.medium_n)[:2] == [16, 8]
    assert m.noise == 0.0 and m.integrator == A.EM and m.neighbors == A.CELL_LINKED
    assert ABM(2, medium={"NMedium": [8, 8]}).model_desc(0.1).medium_alg == A.MEDIUM_FTCS


def test_de_variables_come_first_and_set_user_de():
    law = CBMModels.custom("struct UserLaw {};", [0.5])
    abm = ABM(2, agentODE=law, agent={"a": float, "age": float}, deVariables=("age",))
    assert abm.user_fields == {"age": A.F_U0, "a": A.F_U1}
    assert abm.model_desc(0.1).user_de == 1 and abm.model_desc(0.1).law == A.LAW_CUSTOM
    with pytest.raises(ValueError):
        ABM(2, agent={"age": float}, deVariables=("age",))     # built-in law: no de body
    with pytest.raises(ValueError):
        ABM(2, agentODE=law, agent={"a": float}, deVariables=("b",))
    with pytest.raises(ValueError, match="more than once"):
        ABM(2, agentODE=law, agent={"a": float, "b": float}, deVariables=("a", "a"))


def test_user_parameter_fields_u0_to_u7():
    assert [A.user_field(k) for k in range(8)] == [17, 18, 19, 20, 27, 28, 29, 30]
    with pytest.raises(ValueError):
        A.user_field(8)
    from cbm.api import ABM
    with pytest.raises(ValueError, match="at most 8"):
        ABM(2, agentRule="struct UserRule {};", agent={f"p{k}": float for k in range(9)})
    abm = ABM(2, agentRule="struct UserRule {};", agent={f"p{k}": float for k in range(8)})
    assert list(abm.user_fields.values()) == [17, 18, 19, 20, 27, 28, 29, 30]
