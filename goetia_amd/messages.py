"""Messages of the streaming driver, as goetia.messages defines them
(messages.py:18-120): frozen keyword-only dataclasses with the same names and
fields, serialised to YAML (goetia's mashumaro mixins are not used here; the
field names and values are the same)."""
from dataclasses import asdict, dataclass, field
from enum import Enum, auto
from typing import List

import yaml


class MessageType(Enum):
    Interval = auto()
    SampleStarted = auto()
    SampleFinished = auto()
    SampleSaturated = auto()
    Error = auto()
    DistanceCalc = auto()
    EndStream = auto()


class AllMessages:
    """Marker base class: a listener on AllMessages receives every message."""

    def to_dict(self):
        d = asdict(self)
        d['msg_type'] = self.msg_type.name
        return d

    def to_yaml(self):
        return yaml.safe_dump(self.to_dict(), sort_keys=False)


@dataclass(frozen=True, kw_only=True)
class Interval(AllMessages):
    t: int
    sequence: int
    sample_name: str
    file_names: List[str]
    start_time_seconds: float = 0
    seconds_elapsed_total: float = 0
    seconds_elapsed_sample: float = 0
    seconds_elapsed_interval: float = 0
    modulus: int = 0
    msg_type: MessageType = field(default=MessageType.Interval)


@dataclass(frozen=True, kw_only=True)
class SampleStarted(AllMessages):
    t: int
    sequence: int
    sample_name: str
    file_names: List[str]
    seconds_elapsed_total: float = 0
    msg_type: MessageType = field(default=MessageType.SampleStarted)


@dataclass(frozen=True, kw_only=True)
class SampleFinished(AllMessages):
    t: int
    sequence: int
    sample_name: str
    file_names: List[str]
    seconds_elapsed_total: float = 0
    seconds_elapsed_sample: float = 0
    msg_type: MessageType = field(default=MessageType.SampleFinished)


@dataclass(frozen=True, kw_only=True)
class SampleSaturated(AllMessages):
    t: int
    sequence: int
    sample_name: str
    file_names: List[str]
    msg_type: MessageType = field(default=MessageType.SampleSaturated)


@dataclass(frozen=True, kw_only=True)
class Error(AllMessages):
    t: int
    sequence: int
    sample_name: str
    error: str
    file_names: List[str]
    msg_type: MessageType = field(default=MessageType.Error)


@dataclass(frozen=True, kw_only=True)
class DistanceCalc(AllMessages):
    t: int
    sequence: int
    sample_name: str
    stat_type: str
    distance: float
    stat: float
    file_names: List[str]
    seconds_elapsed_total: float = 0
    seconds_elapsed_sample: float = 0
    seconds_elapsed_interval: float = 0
    msg_type: MessageType = field(default=MessageType.DistanceCalc)


@dataclass(frozen=True, kw_only=True)
class EndStream(AllMessages):
    t: int
    sequence: int
    seconds_elapsed_total: float = 0
    msg_type: MessageType = field(default=MessageType.EndStream)
