"""The streaming driver (AsyncSequenceProcessor, goetia processors.py:133-351)
and its messages (messages.py): per sample SampleStarted, one Interval per
advance, SampleFinished and EndStream, delivered to listeners of both
channels in order; an error ends the run with Error + EndStream; the echo
writes every message as YAML.  Host code (a stand-in processor)."""
import yaml

from goetia_amd.messages import AllMessages, EndStream, Error, Interval, SampleFinished, SampleStarted
from goetia_amd.processors import AsyncSequenceProcessor, every_n_intervals


class Steps:
    """chunked_process yields (n_sequences, time, n_skipped) per advance."""

    def __init__(self, fail_on=None):
        self.fail_on = fail_on
        self.n = 0

    def chunked_process(self, path):
        if path == self.fail_on:
            raise RuntimeError('bad input')
        for _ in range(3):
            self.n += 10
            yield self.n, self.n * 100, 0


def test_messages_in_order(tmp_path):
    samples = [(('a.fq',), 'A'), (('b.fq',), 'B')]
    echo = tmp_path / 'echo.yaml'
    proc = AsyncSequenceProcessor(Steps(), iter(samples), echo=str(echo))
    events, intervals, every2 = [], [], []
    proc.add_listener('events_q', 'all').on_message(AllMessages, events.append)
    proc.add_listener('worker_q', 'iv').on_message(Interval, lambda m: intervals.append((m.sequence, m.t)))

    async def on2(msg):
        every2.append(msg.sequence)

    proc.add_listener('worker_q', 'every2').on_message(Interval, every_n_intervals(on2, 2))
    proc.run()
    kinds = [type(m) for m in events]
    one = [SampleStarted, Interval, Interval, Interval, SampleFinished, EndStream]
    assert kinds == one + one
    assert intervals == [(10, 1000), (20, 2000), (30, 3000), (40, 4000), (50, 5000), (60, 6000)]
    assert every2 == [20, 40, 60]
    assert [m.sample_name for m in events if isinstance(m, SampleStarted)] == ['A', 'B']
    assert events[4].sequence == 30 and events[4].file_names == ['a.fq']
    docs = [yaml.safe_load(block) for block in echo.read_text().split('\n\n') if block.strip()]
    assert [d['msg_type'] for d in docs] == [k.__name__ for k in kinds]
    assert proc.processed == {('a.fq',), ('b.fq',)}


def test_error_ends_the_run():
    samples = [(('a.fq',), 'A'), (('bad.fq',), 'B'), (('c.fq',), 'C')]
    proc = AsyncSequenceProcessor(Steps(fail_on='bad.fq'), iter(samples))
    events = []
    proc.add_listener('events_q', 'all').on_message(AllMessages, events.append)
    proc.run()
    kinds = [type(m).__name__ for m in events]
    assert kinds[-3:] == ['SampleStarted', 'Error', 'EndStream']
    err = [m for m in events if isinstance(m, Error)][0]
    assert 'bad input' in err.error and err.sample_name == 'B'
    assert proc.processed == {('a.fq',)}


def test_messages_are_keyword_only_and_frozen():
    m = Interval(t=1, sequence=2, sample_name='s', file_names=['f'])
    try:
        Interval(1, 2, 's', ['f'])
        raise AssertionError('positional arguments accepted')
    except TypeError:
        pass
    try:
        m.t = 3
        raise AssertionError('message is mutable')
    except Exception as e:
        assert 'cannot assign' in str(e) or 'frozen' in str(e).lower()
    assert m.to_dict()['msg_type'] == 'Interval'
