"""Object caching (reference: pyop2/caching.py:41-96).

``Cached.__new__`` runs ``_process_args`` then ``_cache_key``; a hit returns the
stored object (whose ``__init__`` still runs and must return early when
``_initialized`` is set), a miss constructs and stores it.  A ``None`` key
means "do not cache"."""


class Cached(object):

    def __new__(cls, *args, **kwargs):
        args, kwargs = cls._process_args(*args, **kwargs)
        key = cls._cache_key(*args, **kwargs)
        try:
            return cls._cache_lookup(key)
        except KeyError:
            obj = super(Cached, cls).__new__(cls)
            obj._key = key
            obj._initialized = False
            obj.__init__(*args, **kwargs)
            if key is not None:
                cls._cache_store(key, obj)
            return obj

    @classmethod
    def _cache_lookup(cls, key):
        if key is None:
            raise KeyError(key)
        return cls._cache[key]

    @classmethod
    def _cache_store(cls, key, val):
        cls._cache[key] = val

    @classmethod
    def _process_args(cls, *args, **kwargs):
        return args, kwargs

    @classmethod
    def _cache_key(cls, *args, **kwargs):
        return tuple(args) + tuple((k, v) for k, v in kwargs.items())

    @property
    def cache_key(self):
        """Cache key."""
        return self._key
