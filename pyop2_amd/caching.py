"""Identity cache of objects on another object (semantics of
``pyop2/caching.py:68-150``, ``ObjectCached``): constructing e.g. a second
``Sparsity`` with the same DataSets and maps returns the first one, cached
in the row Set's ``_cache`` dict."""


class ObjectCached:
    """Base class for objects cached on another object.

    Subclasses implement ``_process_args`` (returning ``((cache_obj,) +
    args, kwargs)``) and ``_cache_key`` (``None`` = do not cache).  As in
    the reference, ``__init__`` runs again on a cache hit and must return
    early when ``self._initialized`` is set.
    """

    @classmethod
    def _process_args(cls, *args, **kwargs):
        raise NotImplementedError("Subclass must implement _process_args")

    @classmethod
    def _cache_key(cls, *args, **kwargs):
        raise NotImplementedError("Subclass must implement _cache_key")

    def __new__(cls, *args, **kwargs):
        args, kwargs = cls._process_args(*args, **kwargs)
        cache_obj, args = args[0], args[1:]
        key = cls._cache_key(*args, **kwargs)

        def make():
            obj = super(ObjectCached, cls).__new__(cls)
            obj._initialized = False
            obj.__init__(*args, **kwargs)
            return obj

        if key is None or cache_obj is None:
            return make()
        try:
            cache = cache_obj._cache
        except AttributeError:
            raise RuntimeError("Provided caching object does not have a '_cache' attribute.")
        try:
            return cache[key]
        except KeyError:
            obj = cache[key] = make()
            return obj
