"""The streaming driver over a chunk processor: goetia's AsyncSequenceProcessor
interface (goetia/processors.py:38-351 -- QueueManager, MessageHandler,
RunState, AsyncSequenceProcessor, every_n_intervals; the channel names
'worker_q' and 'events_q'; the message types of messages.py and their order
per sample: SampleStarted, one Interval per processor advance, SampleFinished
or Error, then EndStream), built on asyncio since curio is not a dependency.

Design:
  * A channel (QueueManager) fans every message out to its subscribers
    synchronously on the event loop's thread; its inlet accepts messages from
    any thread (the worker's included) and keeps arrival order.  There is no
    pump task per channel: dispatch() only waits for the channel to close.
  * A listener (MessageHandler) owns an inbox and subscribes when it is
    created, so nothing published after add_listener() can be missed and the
    run needs no start-up delay.
  * The worker thread turns each sample into messages with a generator
    (_sample_messages) that times the processor's advances; the run loop
    posts them and checks the run state between messages.
"""
import asyncio
import collections
import functools
import inspect
import signal
import sys
import threading
import time
import traceback
from enum import Enum, unique

from .messages import AllMessages, EndStream, Error, Interval, SampleFinished, SampleStarted

_CLOSE = None   # end-of-stream marker carried through channels and inboxes


class _Inbox:
    """A FIFO that any thread may put() into and that coroutines read with
    get().  It attaches to the first event loop that reads it (or that a
    channel attaches it to); items put earlier wait in a buffer.  With
    `deliver`, items are handed to that callback on the loop's thread
    instead of being queued."""

    def __init__(self, deliver=None):
        self._deliver = deliver
        self._loop = None
        self._queue = None
        self._pending = collections.deque()
        self._guard = threading.Lock()

    def attach(self, loop):
        with self._guard:
            if self._loop is not None:
                return
            self._loop = loop
            if self._deliver is None:
                self._queue = asyncio.Queue()
            early, self._pending = list(self._pending), collections.deque()
        for item in early:
            self._accept(item)

    def _accept(self, item):   # runs on the loop's thread
        if self._deliver is not None:
            self._deliver(item)
        else:
            self._queue.put_nowait(item)

    def put(self, item):
        with self._guard:
            loop = self._loop
            if loop is None:
                self._pending.append(item)
                return
        try:
            here = asyncio.get_running_loop()
        except RuntimeError:
            here = None
        if here is loop:
            self._accept(item)
        else:
            loop.call_soon_threadsafe(self._accept, item)

    async def get(self):
        self.attach(asyncio.get_running_loop())
        return await self._queue.get()

    def task_done(self):
        """Kept for the queue interface; delivery needs no acknowledgement."""


# goetia's listeners name their queues UniversalQueue
UniversalQueue = _Inbox


class QueueManager:
    """A named broadcast channel (processors.py:38-67): `q` is its inlet,
    subscribers are anything with put() (listener inboxes, other channels'
    inlets)."""

    def __init__(self, q, name):
        self.name = name
        self._subs = collections.OrderedDict()
        self._closed = None
        self.q = q if q is not None else _Inbox()
        self.q._deliver = self._fan_out

    @property
    def subscribers(self):
        return set(self._subs)

    @property
    def subscriber_names(self):
        return dict(self._subs)

    def subscribe(self, q, name):
        self._subs.setdefault(q, name)

    def unsubscribe(self, q):
        self._subs.pop(q, None)

    def attach(self, loop):
        if self._closed is None:
            self._closed = loop.create_future()
        self.q.attach(loop)

    def _fan_out(self, msg):
        for sub in list(self._subs):
            sub.put(msg)
        if msg is _CLOSE and self._closed is not None and not self._closed.done():
            self._closed.set_result(True)

    async def kill(self):
        self.q.put(_CLOSE)

    async def dispatch(self):
        """Completes once the channel has delivered its end-of-stream."""
        self.attach(asyncio.get_running_loop())
        await self._closed


class MessageHandler:
    """Callbacks per message type on one channel (processors.py:70-119);
    callbacks registered for AllMessages see every message after the
    type-specific ones."""

    def __init__(self, name, subscription):
        self.name = name
        self.subscription = subscription
        self.handlers = collections.defaultdict(list)
        self._inbox = _Inbox()
        subscription.subscribe(self._inbox, name)

    def on_message(self, msg_class, callback, *args, **kwargs):
        if not isinstance(msg_class, type):
            raise TypeError('on_message takes a message class')
        self.handlers[msg_class].append((callback, args, kwargs))

    def on_messages(self, msg_classes, callback, *args, **kwargs):
        for cls in msg_classes:
            self.on_message(cls, callback, *args, **kwargs)

    def _callbacks(self, msg):
        return self.handlers.get(type(msg), []) + self.handlers.get(AllMessages, [])

    async def task(self, error_handler=None):
        try:
            while True:
                msg = await self._inbox.get()
                if msg is _CLOSE:
                    return
                for fn, args, kwargs in self._callbacks(msg):
                    out = fn(msg, *args, **kwargs)
                    if inspect.isawaitable(out):
                        await out
        except asyncio.CancelledError:
            raise
        except Exception as e:   # a failing callback ends this listener only
            if error_handler is None:
                raise
            error_handler(e)
        finally:
            self.subscription.unsubscribe(self._inbox)


@unique
class RunState(Enum):
    READY = 0
    RUNNING = 1
    SIGINT = 2
    STOP_SATURATED = 3
    STOP_ERROR = 4
    STOP = 5


class _Stopwatch:
    """Seconds since construction and since the last lap()."""

    def __init__(self):
        self.origin = self.mark = time.perf_counter()

    def total(self):
        return time.perf_counter() - self.origin

    def lap(self):
        now = time.perf_counter()
        start, self.mark = self.mark, now
        return start, now - start


class AsyncSequenceProcessor:
    """Advances a chunk processor (anything with chunked_process(*files)
    yielding (n_sequences, time, n_skipped) per advance, such as
    InserterProcessor or FilterProcessor) through `sample_iter` -- pairs of
    (files, sample name) -- on a worker thread (processors.py:133-351).
    Messages go to the 'worker_q' channel, which feeds 'events_q'."""

    def __init__(self, processor, sample_iter, echo=None, broadcast_socket=None):
        self.processor = processor
        self.sample_iter = sample_iter
        self.worker_q = _Inbox()
        self.events_q = _Inbox()
        self.worker_subs = QueueManager(self.worker_q, 'worker_q')
        self.events_subs = QueueManager(self.events_q, 'events_q')
        self.channels = collections.OrderedDict((c.name, c) for c in (self.worker_subs, self.events_subs))
        self.subscribe('worker_q', self.events_q, 'events_q')
        self.listener_tasks = []
        self.run_echo = echo is not None
        self.echo_file = '/dev/stderr' if echo is True else echo
        self.broadcast_socket = broadcast_socket
        self.state = RunState.READY
        self.processed = set()
        self.seconds_elapsed = 0

    def get_channel(self, channel):
        if channel not in self.channels:
            raise KeyError(f'no channel named {channel!r}; channels are {list(self.channels)}')
        return self.channels[channel]

    def subscribe(self, channel_name, collection_q, subscriber_name):
        self.get_channel(channel_name).subscribe(collection_q, subscriber_name)

    def unsubscribe(self, channel_name, collection_q):
        self.get_channel(channel_name).unsubscribe(collection_q)

    def add_listener(self, channel_name, subscriber_name):
        handler = MessageHandler(subscriber_name, self.get_channel(channel_name))
        self.listener_tasks.append(handler.task)
        return handler

    # ---- the worker thread
    def _sample_messages(self, files, name, run_clock, position):
        """Messages of one sample, in order; `position` is the running
        [n_sequences, time] shared across samples."""
        sample_clock = _Stopwatch()
        common = dict(sample_name=name, file_names=list(files))
        yield SampleStarted(t=position[1], sequence=position[0], seconds_elapsed_total=run_clock.total(), **common)
        advances = None
        while True:
            try:
                if advances is None:
                    advances = iter(self.processor.chunked_process(*files))
                n_seqs, t, _skipped = next(advances)
            except StopIteration:
                break
            except Exception as e:
                yield Error(t=position[1], sequence=position[0], error=''.join(
                    traceback.format_exception(type(e), e, e.__traceback__)), **common)
                return
            position[:] = [n_seqs, t]
            start, spent = sample_clock.lap()
            yield Interval(t=t, sequence=n_seqs, seconds_elapsed_interval=spent,
                           seconds_elapsed_sample=time.perf_counter() - sample_clock.origin,
                           seconds_elapsed_total=run_clock.total(), start_time_seconds=start, **common)
        self.processed.add(tuple(files))
        yield SampleFinished(t=position[1], sequence=position[0],
                             seconds_elapsed_sample=time.perf_counter() - sample_clock.origin,
                             seconds_elapsed_total=run_clock.total(), **common)

    def worker(self):
        run_clock = _Stopwatch()
        position = [0, 0]
        for files, name in self.sample_iter:
            failed = False
            for msg in self._sample_messages(files, name, run_clock, position):
                if isinstance(msg, Interval):
                    if self.state is RunState.STOP_SATURATED:
                        failed = True
                        break
                    if self.state is RunState.SIGINT:
                        msg = Error(t=position[1], sequence=position[0], sample_name=name, file_names=list(files),
                                    error='stopped by SIGINT')
                        failed = True
                self.worker_q.put(msg)
                if isinstance(msg, Error):
                    failed = True
                    break
            self.worker_q.put(EndStream(t=position[1], sequence=position[0],
                                        seconds_elapsed_total=run_clock.total()))
            if failed:
                break
        self.seconds_elapsed = run_clock.total()

    # ---- the event loop side
    def _echo_listener(self):
        listener = self.add_listener('events_q', 'echo')
        path = self.echo_file
        stream = {'/dev/stdout': sys.stdout, '/dev/stderr': sys.stderr}.get(path)
        sink = stream if stream is not None else open(path, 'a')

        def write(msg):
            sink.write(msg.to_yaml() + '\n')
            sink.flush()

        listener.on_message(AllMessages, write)
        return sink if stream is None else None

    async def start(self, extra_tasks=None):
        loop = asyncio.get_running_loop()
        for ch in self.channels.values():
            ch.attach(loop)
        owned = self._echo_listener() if self.run_echo else None
        tasks = [asyncio.ensure_future(ch.dispatch()) for ch in self.channels.values()]
        tasks += [asyncio.ensure_future(t()) for t in list(self.listener_tasks) + list(extra_tasks or ())]
        if threading.current_thread() is threading.main_thread():
            signal.signal(signal.SIGINT, lambda signo, frame: self.interrupt())
        self.state = RunState.RUNNING
        try:
            await loop.run_in_executor(None, self.worker)
        finally:
            await self.worker_subs.kill()
            results = await asyncio.gather(*tasks, return_exceptions=True)
            if owned is not None:
                owned.close()
        for r in results:
            if isinstance(r, BaseException):
                print(f'listener failed: {r!r}', file=sys.stderr)

    def run(self, extra_tasks=None):
        """start() on a new event loop; returns once every message is delivered."""
        asyncio.run(self.start(extra_tasks))

    def stop(self):
        self.state = RunState.STOP

    def interrupt(self):
        self.state = RunState.SIGINT

    def saturate(self):
        self.state = RunState.STOP_SATURATED


def every_n_intervals(func, n=1):
    """Wraps an async Interval callback so it runs on every n-th Interval
    (processors.py:344-351)."""
    seen = [0]

    @functools.wraps(func)
    async def wrapped(msg, *args, **kwargs):
        if not isinstance(msg, Interval):
            raise TypeError('every_n_intervals wraps Interval callbacks')
        seen[0] += 1
        if seen[0] % n == 0:
            await func(msg, *args, **kwargs)
    return wrapped
