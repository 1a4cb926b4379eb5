"""Context.copy_settings replays the launch-tuning setters of one context on another (used to
give a graph-capture context the process-wide tuning); checked with a stand-in library."""
from bosonsupremacy_amd import runtime


class _FakeLib:
    def __init__(self):
        self.log = []

    def __getattr__(self, name):
        def fn(handle, *args):
            self.log.append((name,) + args)
            return 0
        return fn


def _ctx():
    c = runtime.Context.__new__(runtime.Context)
    c.lib = _FakeLib()
    c.handle = None
    c.settings = {}
    return c


def test_copy_settings_keeps_last_value_and_order():
    a = _ctx()
    a.call("bsmps_set_group_lanes", 14, 32)
    a.set_tuning(4, 64)                      # resets every size's lane count
    a.call("bsmps_set_group_lanes", 12, 16)
    a.call("bsmps_set_split", 12, 0)
    a.call("bsmps_set_split", 12, 1)
    a.call("bsmps_set_contract_form", 1, 1)
    a.call("bsmps_synchronize")              # not a setter: not recorded
    b = _ctx()
    b.copy_settings(a)
    assert b.lib.log == [("bsmps_set_tuning", 4, 64), ("bsmps_set_group_lanes", 12, 16),
                         ("bsmps_set_split", 12, 1), ("bsmps_set_contract_form", 1, 1)]
    assert b.settings == a.settings
