"""Interpolation of log det, trace and Schatten norms of A + tB over t from
estimates at a few points (imate/interpolator/__init__.py), the points
estimated on the device by the SLQ path."""

from .interpolate_schatten import InterpolateSchatten
from .interpolate_logdet import InterpolateLogdet
from .interpolate_trace import InterpolateTrace

__all__ = ['InterpolateSchatten', 'InterpolateLogdet', 'InterpolateTrace']
