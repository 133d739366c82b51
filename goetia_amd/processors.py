"""goetia's chunk driver for the dBG path: IntervalCounter (metrics.hh:101-145),
InserterProcessor<dBG> (processors.hh:248-344) on the native inserter
(gt_inserter_*) and FilterProcessor<Filter> (processors.hh:355-440), both over
FileProcessor (processors.hh:45-236).

advance() processes reads until the k-mers of the reads since the last
interval reach `interval`, returning right after that read with the running
(n_sequences, time, remaining) as goetia's does; reads are parsed and
processed in batches on the GPU instead of one call each.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .parsing import FastxParser, fastx_text
from .reads import ReadBatch


class IntervalCounter:
    """goetia::IntervalCounter: poll(incr) adds incr and returns True, resetting
    the counter, once it reaches the interval."""
    DEFAULT_INTERVAL = 500000

    def __init__(self, interval=DEFAULT_INTERVAL):
        self.interval = interval
        self._counter = 0
        self._total = 0

    def counter(self):
        return self._counter

    def total(self):
        return self._total

    def poll(self, incr=1):
        self._counter += incr
        self._total += incr
        if self._counter >= self.interval:
            self._counter = 0
            return True
        return False


class InserterProcessor:
    """goetia::InserterProcessor<dBG<S, Sh>> for one graph."""
    DEFAULT_INTERVAL = IntervalCounter.DEFAULT_INTERVAL

    def __init__(self, graph, interval=DEFAULT_INTERVAL, verbose=False, batch_bases=64 << 20):
        self.inserter = graph
        self.verbose = verbose
        self._interval = interval
        self._h = N.vp()
        N.call('gt_inserter_new', graph.S.handle, graph.K, int(graph.canonical), graph.mode, int(interval),
               int(batch_bases), C.byref(self._h))

    @classmethod
    def build(cls, graph, interval=DEFAULT_INTERVAL, verbose=False):
        return cls(graph, interval, verbose)

    def __del__(self):
        try:
            if self._h.value:
                N.load().gt_inserter_free(self._h)
                self._h = N.vp()
        except Exception:
            pass

    def advance(self, parser):
        """FileProcessor::advance: (n_sequences, time, remaining)."""
        n, t, rem = C.c_uint64(), C.c_uint64(), C.c_int()
        N.call('gt_inserter_advance', self._h, parser._h, C.byref(n), C.byref(t), C.byref(rem))
        return n.value, t.value, bool(rem.value)

    def process(self, filename, strict=False, min_length=0):
        """FileProcessor::process (processors.hh:100-126): the whole file;
        (n_sequences, time)."""
        parser = filename if isinstance(filename, FastxParser) else FastxParser(filename, strict, min_length)
        remaining = True
        n = t = 0
        while remaining:
            n, t, remaining = self.advance(parser)
        return n, t

    def chunked_process(self, file, right_file=None):
        """pythonize_processors.py:12-33: advance through a file (or a
        parser), yielding (n_sequences, time, parser.n_skipped()) after every
        advance that consumed reads."""
        if right_file is not None:
            raise N.UnsupportedError('paired-end input (SplitPairedReader) is not supported')
        parser = FastxParser(file) if isinstance(file, (str, bytes)) else file
        advancing = True
        prev = 0
        while advancing:
            n, t, advancing = self.advance(parser)
            if n > prev:
                yield n, t, parser.n_skipped()
            prev = n

    def _stats(self):
        n, t, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        N.call('gt_inserter_stats', self._h, C.byref(n), C.byref(t), C.byref(c))
        return n.value, t.value, c.value

    def n_sequences(self):
        return self._stats()[0]

    def time_elapsed(self):
        return self._stats()[1]

    def interval(self):
        return self._interval


class FilterProcessor:
    """goetia::FilterProcessor<Filter> (processors.hh:355-440): every read goes
    through the filter (StreamingSolidFilter or DiginormFilter) and the reads
    that pass are written to output_filename with Record::write_fastx
    (parsing.hh:38-49).  A read's time is its k-mer count (0 below K, which
    the reference's filter rejects with SequenceLengthException); advance()
    returns at the read that brings the IntervalCounter to the interval, and
    the filter runs on the GPU over the reads up to there, so the graph at a
    return holds exactly the reads before it."""
    DEFAULT_INTERVAL = IntervalCounter.DEFAULT_INTERVAL

    def __init__(self, filter, output_filename, interval=DEFAULT_INTERVAL, verbose=False, batch_reads=1 << 16,
                 batch_bases=16 << 20):
        self.filter = filter
        self.K = filter.K
        self.verbose = verbose
        self._interval = int(interval)
        self._timer = IntervalCounter(interval)
        self._out = open(output_filename, 'w')
        self._n_sequences = 0
        self._n_passed = 0
        self._pending = None        # [batch, names, qualities, times, next read]
        self._batch_reads, self._batch_bases = batch_reads, batch_bases

    @classmethod
    def build(cls, filter, output_filename, interval=DEFAULT_INTERVAL, verbose=False):
        return cls(filter, output_filename, interval, verbose)

    def close(self):
        if self._out is not None and not self._out.closed:
            self._out.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _run(self, batch, names, quals, i, j):
        o = batch.offsets
        sub = ReadBatch(batch.seq[o[i]:o[j]], o[i:j + 1] - o[i])
        passed = self.filter.filter_sequences(sub)
        keep = np.nonzero(passed)[0]
        if len(keep):
            seq = batch.seq
            self._out.write(''.join(fastx_text(names[i + k], seq[o[i + k]:o[i + k + 1]].tobytes().decode(),
                                               quals[i + k].decode() if quals[i + k] else None) for k in keep))
        self._n_passed += len(keep)

    def advance(self, parser):
        """FileProcessor::advance (processors.hh:200-222):
        (n_sequences, time, remaining)."""
        while True:
            if self._pending is None:
                batch, names, quals = parser.next_batch(self._batch_reads, self._batch_bases, quality=True)
                if batch.n_reads == 0:
                    self._out.flush()
                    return self._n_sequences, self._timer.total(), False
                times = np.maximum(np.diff(batch.offsets) - self.K + 1, 0)
                self._pending = [batch, names, quals, times, 0]
            batch, names, quals, times, i = self._pending
            reach = self._timer.counter() + np.cumsum(times[i:])
            hit = np.nonzero(reach >= self._interval)[0]
            j = i + (int(hit[0]) + 1 if len(hit) else len(times) - i)
            self._run(batch, names, quals, i, j)
            crossed = self._timer.poll(int(times[i:j].sum()))
            self._n_sequences += j - i
            if j >= batch.n_reads:
                self._pending = None
            else:
                self._pending[4] = j
            if crossed:
                self._out.flush()
                return self._n_sequences, self._timer.total(), True

    def process(self, filename, strict=False, min_length=0):
        """FileProcessor::process (processors.hh:100-126): (n_sequences, time)."""
        parser = filename if isinstance(filename, FastxParser) else FastxParser(filename, strict, min_length)
        remaining = True
        n = t = 0
        while remaining:
            n, t, remaining = self.advance(parser)
        return n, t

    def chunked_process(self, file, right_file=None):
        """pythonize_processors.py:12-33."""
        return InserterProcessor.chunked_process(self, file, right_file)

    def n_passed(self):
        return self._n_passed

    def n_sequences(self):
        return self._n_sequences

    def time_elapsed(self):
        return self._timer.total()

    def interval(self):
        return self._interval


# ------------------------------------------------------------ streaming driver
#
# goetia's AsyncSequenceProcessor (processors.py:133-351): a worker thread
# advances a processor through a list of samples and pushes SampleStarted /
# Interval / SampleFinished / Error / EndStream messages onto a worker queue,
# which is forwarded to an events queue; listeners subscribe to either and get
# callbacks per message type.  goetia runs it on curio; this is the same
# object model on asyncio (curio is not a dependency here).

import asyncio  # noqa: E402
import functools  # noqa: E402
import inspect  # noqa: E402
import signal  # noqa: E402
import sys  # noqa: E402
import threading  # noqa: E402
import time  # noqa: E402
import traceback  # noqa: E402
from collections import OrderedDict, defaultdict  # noqa: E402
from enum import Enum, unique  # noqa: E402

from .messages import AllMessages, EndStream, Error, Interval, SampleFinished, SampleStarted  # noqa: E402

DEFAULT_INTERVAL = IntervalCounter.DEFAULT_INTERVAL


def time_iterable(iterable):
    """goetia.utils.time_iterable: (item, start, seconds since the previous item)."""
    start = time.perf_counter()
    for result in iterable:
        elapsed = time.perf_counter() - start
        yield result, start, elapsed
        start = time.perf_counter()


class UniversalQueue:
    """A queue that threads and coroutines write and coroutines read (the role
    of curio.UniversalQueue in goetia); bound to the event loop that first
    reads it, items put before that are kept."""

    def __init__(self):
        self._q = None
        self._loop = None
        self._early = []
        self._lock = threading.Lock()

    def _bind(self):
        with self._lock:
            if self._q is None:
                self._loop = asyncio.get_running_loop()
                self._q = asyncio.Queue()
                for item in self._early:
                    self._q.put_nowait(item)
                self._early = []

    def put(self, item):
        """Never blocks; callable from any thread or coroutine."""
        with self._lock:
            if self._q is None:
                self._early.append(item)
                return
            loop = self._loop
        try:
            running = asyncio.get_running_loop()
        except RuntimeError:
            running = None
        if running is loop:
            self._q.put_nowait(item)
        else:
            loop.call_soon_threadsafe(self._q.put_nowait, item)

    async def get(self):
        self._bind()
        return await self._q.get()

    def task_done(self):
        pass


class QueueManager:
    """processors.py:38-67: a channel forwarding its queue to subscribers."""

    def __init__(self, q, name):
        self.q = q
        self.name = name
        self.subscribers = set()
        self.subscriber_names = {}

    def subscribe(self, q, name):
        if q not in self.subscribers:
            self.subscribers.add(q)
            self.subscriber_names[q] = name

    def unsubscribe(self, q):
        self.subscribers.discard(q)
        self.subscriber_names.pop(q, None)

    async def kill(self):
        self.q.put(None)

    async def dispatch(self):
        while True:
            msg = await self.q.get()
            for sub_q in list(self.subscribers):
                sub_q.put(msg)
            self.q.task_done()
            if msg is None:
                break


class MessageHandler:
    """processors.py:70-119: callbacks per message type on a subscription."""

    def __init__(self, name, subscription):
        self.subscription = subscription
        self.name = name
        self.handlers = defaultdict(list)

    async def task(self, error_handler=None):
        msg_q = UniversalQueue()
        self.subscription.subscribe(msg_q, self.name)
        try:
            while True:
                msg = await msg_q.get()
                if msg is None:
                    break
                for callback, args, kwargs in self.handlers[type(msg)] + self.handlers[AllMessages]:
                    if inspect.iscoroutinefunction(callback):
                        await callback(msg, *args, **kwargs)
                    else:
                        callback(msg, *args, **kwargs)
        except asyncio.CancelledError:
            raise
        except Exception as e:
            if error_handler is None:
                raise
            error_handler(e)
        finally:
            self.subscription.unsubscribe(msg_q)

    def on_message(self, msg_class, callback, *args, **kwargs):
        assert type(msg_class) is type
        self.handlers[msg_class].append((callback, args, kwargs))

    def on_messages(self, msg_classes, callback, *args, **kwargs):
        for msg_class in msg_classes:
            self.handlers[msg_class].append((callback, args, kwargs))


@unique
class RunState(Enum):
    READY = 0
    RUNNING = 1
    SIGINT = 2
    STOP_SATURATED = 3
    STOP_ERROR = 4
    STOP = 5


class AsyncSequenceProcessor:
    """processors.py:133-351: manages advancing a processor (anything with
    chunked_process, e.g. InserterProcessor or FilterProcessor) through
    `sample_iter` -- ((file, ...), name) pairs -- on a worker thread; messages
    go to the 'worker_q' channel and from there to 'events_q'."""

    def __init__(self, processor, sample_iter, echo=None, broadcast_socket=None):
        self.worker_q = UniversalQueue()
        self.worker_subs = QueueManager(self.worker_q, 'worker_q')
        self.events_q = UniversalQueue()
        self.events_subs = QueueManager(self.events_q, 'events_q')
        self.channels = OrderedDict()
        self.channels[self.worker_subs.name] = self.worker_subs
        self.channels[self.events_subs.name] = self.events_subs
        self.subscribe('worker_q', self.events_q, 'events_q')
        self.listener_tasks = []
        self.processor = processor
        self.sample_iter = sample_iter
        self.run_echo = echo is not None
        self.echo_file = '/dev/stderr' if echo is True else echo
        self.state = RunState.READY
        self.processed = set()
        self.seconds_elapsed = 0

    def get_channel(self, channel):
        try:
            return self.channels[channel]
        except KeyError:
            print(f'Requested invalid channel: "{channel}" does not exist.', file=sys.stderr)
            raise

    def subscribe(self, channel_name, collection_q, subscriber_name):
        self.get_channel(channel_name).subscribe(collection_q, subscriber_name)

    def unsubscribe(self, channel_name, collection_q):
        self.get_channel(channel_name).unsubscribe(collection_q)

    def add_listener(self, channel_name, subscriber_name):
        listener = MessageHandler(subscriber_name, self.get_channel(channel_name))
        self.listener_tasks.append(listener.task)
        return listener

    def worker(self):
        stream_time, n_seqs = 0, 0
        worker_start_time = time.perf_counter()
        for (sample, name), sample_start_time, _ in time_iterable(self.sample_iter):
            self.worker_q.put(SampleStarted(sample_name=name, file_names=list(sample), t=stream_time,
                                            seconds_elapsed_total=time.perf_counter() - worker_start_time,
                                            sequence=n_seqs))
            try:
                for (n_seqs, stream_time, n_skipped), interval_start_time, interval_elapsed_time in \
                        time_iterable(self.processor.chunked_process(*sample)):
                    if self.state is RunState.STOP_SATURATED:
                        return
                    if self.state is RunState.SIGINT:
                        self.worker_q.put(Error(t=stream_time, sequence=n_seqs, sample_name=name,
                                                file_names=list(sample), error='Process terminated (SIGINT).'))
                        return
                    now = time.perf_counter()
                    self.worker_q.put(Interval(t=stream_time, sequence=n_seqs, sample_name=name,
                                               seconds_elapsed_interval=interval_elapsed_time,
                                               seconds_elapsed_sample=now - sample_start_time,
                                               seconds_elapsed_total=now - worker_start_time,
                                               start_time_seconds=interval_start_time, file_names=list(sample)))
                self.processed.add(tuple(sample))
                now = time.perf_counter()
                self.worker_q.put(SampleFinished(t=stream_time, sequence=n_seqs, sample_name=name,
                                                 seconds_elapsed_sample=now - sample_start_time,
                                                 seconds_elapsed_total=now - worker_start_time,
                                                 file_names=list(sample)))
            except Exception as e:
                self.worker_q.put(Error(t=stream_time, sequence=n_seqs, sample_name=name, file_names=list(sample),
                                        error=''.join(traceback.format_tb(e.__traceback__)) + repr(e)))
                return
            finally:
                self.worker_q.put(EndStream(t=stream_time,
                                            seconds_elapsed_total=time.perf_counter() - worker_start_time,
                                            sequence=n_seqs))

    async def start(self, extra_tasks=None):
        tasks = []
        try:
            for channel in self.channels.values():
                tasks.append(asyncio.create_task(channel.dispatch()))
            if self.run_echo:
                listener = self.add_listener('events_q', 'echo')
                echo_file = self.echo_file

                def echo(msg):
                    mode = 'w' if echo_file in ('/dev/stdout', '/dev/stderr') else 'a'
                    with open(echo_file, mode) as fp:
                        fp.write(f'{msg.to_yaml()}\n')

                listener.on_message(AllMessages, echo)
            for task in self.listener_tasks:
                tasks.append(asyncio.create_task(task()))
            for task in extra_tasks or ():
                tasks.append(asyncio.create_task(task()))
            self.state = RunState.RUNNING
            if threading.current_thread() is threading.main_thread():
                signal.signal(signal.SIGINT, lambda signo, frame: self.interrupt())
            await asyncio.sleep(0.05)
            await asyncio.get_running_loop().run_in_executor(None, self.worker)
            await asyncio.sleep(0.05)
            await self.worker_subs.kill()
            await asyncio.gather(*tasks)
        except Exception as e:
            print(e, file=sys.stderr)

    def run(self, extra_tasks=None):
        """start() on a new event loop, until every message is delivered."""
        asyncio.run(self.start(extra_tasks))

    def stop(self):
        self.state = RunState.STOP

    def interrupt(self):
        self.state = RunState.SIGINT

    def saturate(self):
        self.state = RunState.STOP_SATURATED


def every_n_intervals(func, n=1):
    """processors.py:344-351: call an async Interval handler every n intervals."""
    poller = IntervalCounter(n)

    @functools.wraps(func)
    async def wrapped(msg, *args, **kwargs):
        assert isinstance(msg, Interval)
        if poller.poll():
            await func(msg, *args, **kwargs)
    return wrapped
